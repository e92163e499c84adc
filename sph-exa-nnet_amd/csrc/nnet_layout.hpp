// Kernel configuration and LDS layout of the solve kernels, plus the index functions the kernels
// use to address their LDS arrays.  Everything here is constexpr and free of HIP types, so the
// same definitions compile into the device code and into host-side checks of every LDS index
// the kernels form (clang treats constexpr functions as host and device functions).
#pragma once

namespace nnet_amd
{
constexpr int MAXQ  = 3;  // reactant entries per reaction
constexpr int MAXSP = 88; // species + 1 of the largest network

// The exp / pow values of the six hand-fitted rate groups, evaluated one per lane of a wavefront:
// level 1, NE1 exps and NP1 pows; level 2, NE2 exps of level-1 results.  They pass through the
// Jacobian array sJv (N1 + NE2 entries), which is rewritten only after the rate phase.
namespace spx
{
constexpr int NE1 = 36, NP1 = 13, N1 = NE1 + NP1, NE2 = 6;
}

// Rate slots written by the six hand-fitted rate groups (3-alpha forward/reverse, 12C+12C,
// 12C+16O, 16O+16O, 12C(a,g) forward/reverse, 16O(a,g) forward/reverse).
struct SpecialIdx
{
    int r3a_e, r3a_l, r24, r1216, r32, rcag_e, rcag_l, roag_e, roag_l;
};

// ---------------------------------------------------------------------------------------------
// Kernel configuration per network
// ---------------------------------------------------------------------------------------------
// NT threads per workgroup (one particle), TT x TT tiles of 16 on a WR x WR wave grid, NNZ / IMG
// capacities of the Jacobian pattern and of the LDS assembly image, WPE = waves per SIMD the
// register budget is sized for (net14: 3 measured faster than 4 (spills) and 2 (fewer particles))
template<int NET>
struct Cfg;
template<>
struct Cfg<14>
{
    static constexpr int D = 14, N = 15, NR = 29, NE = 16, NT = 64, TT = 1, WR = 1, NNZ = 72, IMG = 576, WPE = 3;
};
template<>
struct Cfg<86>
{
    static constexpr int D = 86, N = 87, NR = 309, NE = 157, NT = 256, TT = 6, WR = 2, NNZ = 800, IMG = 5248, WPE = 2;
};
template<>
struct Cfg<87>
{
    static constexpr int D = 87, N = 88, NR = 311, NE = 157, NT = 256, TT = 6, WR = 2, NNZ = 800, IMG = 5248, WPE = 2;
};

// LDS of one workgroup, in doubles (the assembly image is a separate array of IMG words)
template<int NET, bool MIX>
struct Lds
{
    using C = Cfg<NET>;
    static constexpr int D = C::D, N = C::N, NR = C::NR, NE = C::NE;
    static constexpr int NU = N * (N + 1) / 2; // packed normalised rows (columns k+1 .. N)
    static constexpr int NL = N * (N - 1) / 2; // packed multipliers (columns 0 .. k-1)
    // head: rate and mass-action arrays, dead once the Jacobian is summed
    static constexpr int oMu = 0, oEff = oMu + D, oDeff = oEff + NE, oL = oDeff + NE, oDl = oL + NE,
                         oRate = oDl + NE, oDrate = oRate + NR, oVal = oDrate + NR, oDval = oVal + NR,
                         oTr = oDval + NR, oJok = oTr + NR * MAXQ, headEnd = oJok + NR;
    // FP32 factors of the mixed path (U rows, multipliers, pivot reciprocals), in doubles
    static constexpr int f32Words = MIX ? (NU + NL + N + 1) / 2 : 0;
    // tail: what the matrix fill and the FP64 refinement residual read
    static constexpr int oBE = (headEnd > f32Words ? headEnd : f32Words), oDYdot = oBE + D, oDYdT = oDYdot + D,
                         oCorr = oDYdT + D, oEv = oCorr + D, oEos = oEv + 12, oJv = oEos + 4,
                         phaseEnd = oJv + C::NNZ;
    // the FP64 packed rows alias everything above once the matrix sits in registers
    static constexpr int unionSize = phaseEnd > NU ? phaseEnd : NU;
    static constexpr int oY = unionSize, oYb = oY + D, oX = oYb + D, oRow0 = oX + N + 1;
    // LU panel buffers (two generations of: panel columns [row][4], pivot rows [column][4]); after
    // the LU they hold the refinement residual
    static constexpr int NPAD = 16 * C::TT, oPan = ((oRow0 + (MIX ? N + 1 : 0)) + 3) & ~3,
                         total = oPan + 16 * NPAD;
};

// Jacobian non-zero slot (position in the image's descending-length order) that lane `tid` of a
// particle's NT lanes sums in pass `pass`, or a value >= nnz when it has none: pass p takes slot
// tid + NT p; with four wavefronts the slots beyond three full passes go to wavefront 1 (no species
// lanes, the lightest load of that phase) rather than to wavefront 0, which holds the longest lists.
constexpr int jac_slot(int tid, int pass, int NT, int nnz)
{
    int k = tid + NT * pass;
    if (NT == 256 && pass >= 3) k = (tid >> 6) == 1 ? 3 * NT + 64 * (pass - 3) + (tid - 64) : nnz;
    return k;
}

// Matrix row held by register v of lane (lr = (l & 63) >> 4, lc = l & 15) in tile row I of the
// register-resident Newton matrix, and its column in tile column J (the v_mfma_f64_16x16x4f64
// result layout).
constexpr int tile_row(int I, int lr, int v) { return 16 * I + lr + 4 * v; }
constexpr int tile_col(int J, int lc) { return 16 * J + lc; }

// Index functions of the tile LU (blocks of four pivots) and of its triangular solves.
template<int NET>
struct LuIdx
{
    using C = Cfg<NET>;
    static constexpr int N = C::N, TT = C::TT, WR = C::WR, TI = TT / WR, NPAD = 16 * TT;
    static constexpr int NS = (N + 3) / 4; // blocks of four pivots

    // normalised upper entry (k, j), k < j <= N: net86 / net87 pack it by columns (column j holds
    // rows 0 .. j-1, so the back substitution's lanes read consecutive words: measured faster),
    // net14 by rows (row k holds columns k+1 .. N: its single wavefront's LU stores are then the
    // consecutive ones, measured faster there); first packed entry of multiplier row k (columns
    // 0 .. k-1)
    static constexpr int uidx(int k, int j)
    {
        return NET == 14 ? k * N - k * (k - 1) / 2 - k - 1 + j : j * (j - 1) / 2 + k;
    }
    static constexpr int lrow(int k) { return k * (k - 1) / 2; }
    // panel buffers of block `step`: generation (step & 1) of [row][4] panel columns P, followed
    // by [column][4] pivot rows RT
    static constexpr int pan(int step) { return (step & 1) * 8 * NPAD; }
    static constexpr int pcol(int step, int row, int q) { return pan(step) + row * 4 + q; }
    static constexpr int prow(int step, int col, int q) { return pan(step) + 4 * NPAD + col * 4 + q; }
    static constexpr int trow(int I, int lr, int v) { return tile_row(I, lr, v); }
    static constexpr int tcol(int J, int lc) { return tile_col(J, lc); }
    // A-operand row m of a tile: FP64 MFMA rows come in result order; the FP32 MFMA returns row
    // 4(l>>4) + v in register v, so with the FP64 placement A-operand row m is (m>>2) + 4(m&3)
    static constexpr int arow(bool fp32, int m) { return fp32 ? (m >> 2) + 4 * (m & 3) : m; }
};
} // namespace nnet_amd
