// libnnet_mi355x: MI355X (gfx950) implementation of SPH-EXA-nnet's batched implicit nuclear
// network solve behind the C-ABI of include/nnet_mi355x.h.
//
// Execution model (FP64 path): the lanes of one particle group (a workgroup of four wavefronts for
// net86 / net87, one wavefront for net14) share the work of one Newton iteration:
//   rates      one lane per fitted rate / hand-fitted group / species (nnet_dev_rates.hpp)
//   assembly   Jacobian entries summed from host-precomputed ordered contribution lists (the
//              values the reference's scatter loops, nuclear_net.hpp:347-438, produce), then the
//              (d+1)x(d+2) augmented Newton matrix is filled directly into registers as 16x16
//              tiles in the v_mfma_f64_16x16x4f64 result layout.
//   LU         right-looking unpivoted elimination on the register-resident tiles, blocks of
//              four pivots, rank-4 trailing updates on MFMA (eigen.hpp:179-206); the normalised
//              pivot rows are staged in LDS, which also serves the back substitution.
//   controller the substep/Newton state machine of nuclear_net.hpp:692-959 is replicated in
//              every lane (uniform control flow) with exact replay-stall detection.
// Persistent workgroups (as many as are resident at once) take particles from a device work queue,
// longest first when the caller passes a cost hint, so particles of very different stiffness
// share the CUs without a tail of idle workgroups.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nnet_mi355x.h"
#include "nnet_dev_rates.hpp"
#include "nnet_dev_eos.hpp"
#include "nnet_tables.h"
#include "nnet_layout.hpp"
#include "nnet_image.hpp"

namespace nnet_amd
{
// nuclear_net.hpp:48-106
namespace k
{
constexpr double theta = 0.8, minTemp = 1e8, minRho = 1e5, maxDtStep = 2, minDtStep = 1e-2, nanDtStep = 2e-1;
constexpr double epsilonSystem = 1e-40, epsilonVector = 1e-16;
constexpr double NR_maxDt = 1e-2, NR_dTOverTempTarget = 1e-2, NR_dTOverTempTol = 4, NR_itTol = 1e-7;
constexpr int    NR_maxIt      = 11;
constexpr double substep_dtTol = 1e-6;
// mixed-precision solve: at most mixed_sweeps(net) FP64-residual refinement sweeps of the FP32
// factors, converged when every correction is below mixedTol of its component's scale.  On net86 /
// net87 an iteration that needs a second sweep mostly fails the third too, so one sweep and an
// early hand-over to the FP64 pass is faster (net86 Helmholtz, 2M particles: 885k against 825k
// particle-steps/s with three); net14's 15x15 systems refine within three sweeps.
constexpr int mixed_sweeps(int net) { return net == 14 ? 3 : 1; }
constexpr double mixedTol    = 1e-11;
} // namespace k

// ---------------------------------------------------------------------------------------------
// Per-network constant data in device memory (built on the host by build_network()).
// ---------------------------------------------------------------------------------------------
struct NetDev
{
    int        net, d, nR, nE, nfit, fit_e0, rev_div, nnz;
    SpecialIdx sp;
    // rate parameterization
    const double* BE0;    // [d'] constant binding energies (erg/mol), d' = 14 or 86
    const double* Z53;    // [d'] Z^(5/3) (host glibc pow: the value the reference computes)
    const double* fit;    // [nfit][8]
    const double* q;      // [nfit]
    const double* choose; // [rows][24]
    const int*    ca;     // [nfit] choose row of the reverse numerator
    const int*    cb;     // [nfit] choose row of the reverse denominator
    const int*    photo;  // [nfit] reverse is a photodisintegration (carries T^3/2)
    const int*    scr;    // [nE][4]: mu index a, b, c, e (e < 0: none)
    const int*    swap;   // [nE]
    const int*    scr_l;  // [nE] reverse rate is screened too
    const int*    push_e; // [nE] rate index of eff[i] (-1: not in the list)
    const int*    push_l; // [nE] rate index of l[i]
    const double* elogT;  // net87 electron-capture table
    const double* elogR;
    const double* etab;
    double        BEe; // net87 electron binding energy (erg/mol)
    // Assembly image, copied into LDS once per workgroup (32-bit words; sub-arrays at the
    // offsets below, in units of their element size):
    //   rx[nR+1]      word: reactant_begin | product_begin << 16 (rx[nR]: entry count)
    //   ent16[]       halfword: species | count << 8 (reference entry order)
    //   sp_ptr16[d+1] halfword: species s sums entries [sp_ptr[s], sp_ptr[s+1]) of sp_ent
    //   sp_ent[]      word: reaction | (signed count & 0xff) << 16, in reaction order
    //   nz_ptr16[nnz+1], nz_ent[]: Jacobian non-zero e (row-major) sums (reaction, reactant
    //                 position, signed count): r | q << 12 | (c & 0xff) << 16, in the order
    //                 of the reference's scatter loops
    //   row_ptr16[d+1], nz_col8[nnz]: non-zeros of row a are [row_ptr[a], row_ptr[a+1]), columns
    //                 ascending
    const uint32_t* img;
    int             img_words, oRx, oEnt16, oSpPtr16, oSpEnt, oNzPtr16, oNzEnt, oRowPtr16, oRowEnt, oJord16, oMaPerm16;
    // per lane, per register slot: index of the Jacobian non-zero the slot holds (-1: zero),
    // laid out [slot][lane] for coalesced loads
    const uint4*   lanemap; // [lane]: register-slot mask (x, y), first non-zero id (z)
};

// Passed by value in the kernel arguments: no device allocation per call.  The single-batch
// kernel takes EosDev / BatchDev with the charges and abundance pointers inline; the fused
// multi-network kernel keeps them in shared arrays and hands the solve an EosView / BatchView.
struct EosCore
{
    int             kind;
    int             zdim;
    double          mu;
    const HelmGrid* grid;
    const double*   cells;
};
struct EosDev : EosCore
{
    double Z[MAXSP];
};
struct EosView : EosCore
{
    const double* Z;
};

struct BatchCore
{
    size_t        n;
    const double* rho;
    double*       rho_m1;
    double*       temp;
    double*       dt;
    uint8_t*      status;
    uint32_t*     substeps;
    uint32_t*     newton;
    uint32_t*     fp64_solves;
    const uint32_t* order; // dispatch order of the particles (nullptr: index order)
    // positions [qrange[0], qrange[1]) of `order` form this launch's work (nullptr: [0, n)); the
    // bounds are device values written by earlier kernels of the same stream
    const unsigned int* qrange;
    unsigned int*       queue; // work-queue counter (zeroed by nnet_queue_reset before the launch)
    // mixed precision: the mixed pass appends the particles whose FP32 refinement failed to
    // redo[0 .. *redo_count); an FP64 pass then integrates them from the start of the step
    uint32_t*     redo;
    unsigned int* redo_count;
    int           fp64_pass; // FP64 pass of a mixed batch: fp64_solves = newton
};
struct BatchDev : BatchCore
{
    double* Y[MAXSP]; // per-species abundance arrays (SoA)
};
struct BatchView : BatchCore
{
    double* const* Y;
};

struct StepParams
{
    double hydro_dt, previous_dt;
    int    use_drhodt;
    long   max_iter;
};

// x^n for the stoichiometric counts 0..3 (std::pow(x, n) with n integral):
// n = 3 is the correctly rounded cube (double-double product), as glibc's pow returns.
__device__ __forceinline__ double pow_count(double x, int n)
{
    if (n == 0) return 1.;
    if (n == 1) return x;
    if (n == 2) return x * x;
    const double p  = x * x;
    const double ep = fma(x, x, -p);
    const double c  = p * x;
    const double ec = fma(p, x, -c);
    return c + (ec + ep * x);
}
__device__ __forceinline__ double factorial_count(int n) { return n == 2 ? 2. : n == 3 ? 6. : n == 4 ? 24. : 1.; }

// ---------------------------------------------------------------------------------------------
// Tile LU.  The augmented (N x N+1) Newton matrix lives in registers as 16x16 FP64 tiles in the
// result layout of v_mfma_f64_16x16x4f64: register v of lane l of a tile holds row 16*I + (l>>4)
// + 4v, column 16*J + (l&15).  Wave (wr, wc) of the WR x WR wave grid owns tiles (wr + WR*ti,
// wc + WR*tj).  Elimination proceeds in blocks of four pivots (eigen.hpp:182-206 order per
// element): the owners publish the block's four columns and four rows to LDS; every lane factors
// the 4x4 diagonal block; one lane per trailing row turns its four panel entries into the
// multipliers (with the reference's 1e-40 skip test: a skipped multiplier is zero, so its update
// adds zero) and one lane per trailing column its four pivot-row entries into the normalised
// ones, in place; after a barrier every wave reads its A/B operands and applies the rank-4
// trailing update to its tiles with one MFMA per tile.  Two barriers per block of four pivots.
// ---------------------------------------------------------------------------------------------
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float  fx4 __attribute__((ext_vector_type(4)));

// value of x in lane `lane` of the wavefront (lane is wave-uniform)
__device__ __forceinline__ double bcast_lane(double x, int lane)
{
    const unsigned long long b  = __double_as_longlong(x);
    const unsigned           lo = __builtin_amdgcn_readlane((unsigned)b, lane);
    const unsigned           hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float bcast_lane(float x, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// Barrier of the lanes that integrate one particle: the workgroup barrier when the particle has
// the workgroup to itself; a wavefront barrier when a one-wavefront particle shares its workgroup
// with others (the LDS instructions of one wavefront execute in program order; the fences keep
// the compiler from moving memory accesses across the barrier).
template<bool WAVE>
__device__ __forceinline__ void psync()
{
    if constexpr (WAVE)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    else
        __syncthreads();
}
// the barrier, returning whether x is true in any lane of the particle
template<bool WAVE>
__device__ __forceinline__ bool psync_or(int x)
{
    if constexpr (WAVE)
    {
        psync<true>();
        return __ballot(x) != 0;
    }
    else
        return __syncthreads_or(x) != 0;
}

template<typename R>
struct TileT;
template<>
struct TileT<double>
{
    using V = dx4;
    static __device__ __forceinline__ V   mfma(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // 1/d from the hardware estimate and two Newton steps (the pivot reciprocal is on the
    // factorization's critical path; the IEEE division sequence is twice as long)
    static __device__ __forceinline__ double recip(double d)
    {
        double r = __builtin_amdgcn_rcp(d);
        r        = fma(r, fma(-d, r, 1.), r);
        return fma(r, fma(-d, r, 1.), r);
    }
};
template<>
struct TileT<float>
{
    using V = fx4;
    static __device__ __forceinline__ V   mfma(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float recip(float d)
    {
        const float r = __builtin_amdgcn_rcpf(d);
        return fmaf(r, fmaf(-d, r, 1.f), r);
    }
};

template<int NET, typename R, bool STORE_L, bool WAVE>
struct TileLU
{
    using C = Cfg<NET>;
    using V = typename TileT<R>::V;
    static constexpr int N = C::N, TT = C::TT, WR = C::WR, TI = TT / WR, NPAD = 16 * TT;
    static constexpr int NS = (N + 3) / 4; // blocks of four pivots
    static_assert(TT % WR == 0 && WR * WR * 64 == C::NT && 16 * TT >= N + 1, "tile grid");

    using X = LuIdx<NET>;
    static_assert(X::NS == NS && X::NPAD == NPAD, "index functions of the same configuration");
    static __device__ __forceinline__ int uidx(int k, int j) { return X::uidx(k, j); }
    static __device__ __forceinline__ int lrow(int k) { return X::lrow(k); }

    static __device__ __forceinline__ R pick4(const R (&x)[4], int q)
    {
        return q == 0 ? x[0] : q == 1 ? x[1] : q == 2 ? x[2] : x[3];
    }

    // the owners of block STEP's columns and rows write them to the panel buffers
    template<int STEP>
    static __device__ __forceinline__ void publish(const V (&Mt)[TI][TI], R* pan, int wr, int wc, int lr, int lc)
    {
        constexpr int p0 = 4 * STEP, Tp = p0 >> 4, c0 = p0 & 15, v0 = c0 >> 2;
        // panel columns [row][q] (columns p0 .. p0+3) and pivot rows [column][q] (rows p0 .. p0+3)
        if (wc == Tp % WR && lc >= c0 && lc < c0 + 4)
        {
            constexpr int tj = Tp / WR;
#pragma unroll
            for (int ti = 0; ti < TI; ++ti)
                if (wr + WR * ti >= Tp)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        pan[X::pcol(STEP, X::trow(wr + WR * ti, lr, v), lc - c0)] = Mt[ti][tj][v];
        }
        if (wr == Tp % WR)
        {
            constexpr int ti = Tp / WR;
#pragma unroll
            for (int tj = 0; tj < TI; ++tj)
                if (wc + WR * tj >= Tp) pan[X::prow(STEP, X::tcol(wc + WR * tj, lc), lr)] = Mt[ti][tj][v0];
        }
    }

    template<int STEP>
    static __device__ __forceinline__ void step(V (&Mt)[TI][TI], R* pan, R* sU, R* sL, R* sRinv, int wr, int wc,
                                                int lr, int lc, int tid)
    {
        constexpr int p0 = 4 * STEP, Tp = p0 >> 4, nq = (N - p0) < 4 ? (N - p0) : 4;
        constexpr int NT = C::NT, HALF = NT / 2;
        static_assert(HALF >= N - 3, "one lane per trailing row and per trailing column");
        const R eps = (R)k::epsilonSystem;
        // opaque copies of the lane coordinates: the block's lane-dependent LDS addresses are
        // formed here, not hoisted to the kernel entry for all 22 blocks at once (which keeps
        // dozens of address registers live across the factorization and spills them); the
        // four-block net14 factorization does not need it
        if constexpr (NET != 14) asm volatile("" : "+v"(tid), "+v"(lr), "+v"(lc));
        auto P  = [&](int row, int q) __attribute__((always_inline)) -> R& { return pan[X::pcol(STEP, row, q)]; };
        auto RT = [&](int col, int q) __attribute__((always_inline)) -> R& { return pan[X::prow(STEP, col, q)]; };

        // 4x4 diagonal block (columns p0 .. p0+3 may include the right-hand side in the last
        // block), factored in every lane
        R Dg[4][4], Un[4][4], Wt[4][4], rinv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                Dg[r][c] = P(p0 + r, c);
                Un[r][c] = 0;
                Wt[r][c] = 0;
            }
#pragma unroll
        for (int q = 0; q < nq; ++q)
        {
            rinv[q] = TileT<R>::recip(Dg[q][q]);
#pragma unroll
            for (int c = q + 1; c < 4; ++c)
                Un[q][c] = Dg[q][c] * rinv[q];
#pragma unroll
            for (int r = q + 1; r < nq; ++r)
            {
                const R w = Dg[r][q];
                Wt[r][q]  = fabs(w) > eps ? w : R(0);
                if (fabs(w) > eps)
#pragma unroll
                    for (int c = q + 1; c < 4; ++c)
                        Dg[r][c] = fma(-w, Un[q][c], Dg[r][c]);
            }
        }
        if (tid == 0)
#pragma unroll
            for (int q = 0; q < nq; ++q)
            {
#pragma unroll
                for (int c = q + 1; c < 4; ++c)
                    if (p0 + c <= N) sU[uidx(p0 + q, p0 + c)] = Un[q][c];
                if constexpr (STORE_L)
                {
                    sRinv[p0 + q] = rinv[q];
#pragma unroll
                    for (int c = 0; c < q; ++c)
                        sL[lrow(p0 + q) + p0 + c] = Wt[q][c];
                }
            }
        // the last block has no trailing rows; net86 and net14 have no trailing columns there
        // either, so the diagonal block's stores above are all it leaves (wavefront 0 runs the
        // triangular solves after the LU and reads them in program order)
        constexpr bool ROWS = p0 + 4 < N, COLS = p0 + 4 <= N;
        if constexpr (!COLS) return;

        if (tid < HALF)
        {
            // one lane per trailing row: its four multipliers, in place of its panel entries
            const int i = p0 + 4 + tid;
            if (i < N)
            {
                R pr[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    pr[c] = P(i, c);
                R Lq[4] = {0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < nq; ++q)
                {
                    const R w = fabs(pr[q]) > eps ? pr[q] : R(0);
                    Lq[q]     = w;
#pragma unroll
                    for (int c = q + 1; c < nq; ++c)
                        pr[c] = fma(-w, Un[q][c], pr[c]);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    P(i, q) = Lq[q];
                if constexpr (STORE_L)
#pragma unroll
                    for (int q = 0; q < nq; ++q)
                        sL[lrow(i) + p0 + q] = Lq[q];
            }
        }
        else
        {
            // one lane per trailing column: the normalised pivot-row entries, in place
            const int j = p0 + 4 + (tid - HALF);
            if (j <= N)
            {
                R u[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    u[q] = RT(j, q);
#pragma unroll
                for (int q = 0; q < nq; ++q)
                {
                    u[q] *= rinv[q];
#pragma unroll
                    for (int r = q + 1; r < nq; ++r)
                        u[r] = fma(-Wt[r][q], u[q], u[r]);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    RT(j, q) = q < nq ? u[q] : R(0);
                    if (q < nq) sU[uidx(p0 + q, j)] = u[q];
                }
            }
        }
        psync<WAVE>();
        // no trailing rows (net87's last block: only the right-hand-side column, normalised by the
        // column lanes above and ordered before the triangular solves by that barrier): no update
        if constexpr (!ROWS) return;

        // rank-4 trailing update, one MFMA per tile that holds trailing rows and columns
        // (A operand: row arow(l&15) of the tile, pivot l>>4; B operand: pivot l>>4, column l&15)
        R Aop[TI], Bop[TI];
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
        {
            const int i = 16 * (wr + WR * ti) + X::arow(std::is_same<R, float>::value, lc);
            Aop[ti]     = (i >= p0 + 4 && i < N) ? -P(i, lr) : R(0);
        }
#pragma unroll
        for (int tj = 0; tj < TI; ++tj)
        {
            const int j = X::tcol(wc + WR * tj, lc);
            Bop[tj]     = (j >= p0 + 4 && j <= N) ? RT(j, lr) : R(0);
        }
        // The tiles holding the next block's panel (its tile row and tile column) are updated
        // first and published; the other tiles' MFMAs are issued after the publish, so they
        // execute while the wave waits at the barrier and factors the next diagonal block (each
        // tile still receives its updates in block order).  A net86 Helmholtz step of 2,097,152
        // particles took 2.35 s against 2.39 s with the publish after all updates.
        constexpr int Tn = (4 * (STEP + 1)) >> 4;
        auto update = [&](bool panel_tiles) __attribute__((always_inline))
        {
#pragma unroll
            for (int ti = 0; ti < TI; ++ti)
#pragma unroll
                for (int tj = 0; tj < TI; ++tj)
                {
                    const int I = wr + WR * ti, J = wc + WR * tj;
                    if (I >= Tp && J >= Tp && 16 * I + 15 >= p0 + 4 && 16 * J + 15 >= p0 + 4 &&
                        ((I == Tn || J == Tn) == panel_tiles))
                        Mt[ti][tj] = TileT<R>::mfma(Aop[ti], Bop[tj], Mt[ti][tj]);
                }
        };
        update(true);
        if constexpr (STEP + 1 < NS) publish<STEP + 1>(Mt, pan, wr, wc, lr, lc);
        __builtin_amdgcn_sched_barrier(0);
        update(false);
        psync<WAVE>();
        if constexpr (STEP + 1 < NS) step<STEP + 1>(Mt, pan, sU, sL, sRinv, wr, wc, lr, lc, tid);
    }

    // forward elimination of the augmented matrix; leaves the normalised rows in sU (and, with
    // STORE_L, the multipliers in sL and the pivot reciprocals in sRinv)
    static __device__ __forceinline__ void factor(V (&Mt)[TI][TI], R* pan, R* sU, R* sL, R* sRinv, int wr, int wc,
                                                  int lr, int lc, int tid)
    {
        publish<0>(Mt, pan, wr, wc, lr, lc);
        psync<WAVE>();
        step<0>(Mt, pan, sU, sL, sRinv, wr, wc, lr, lc, tid);
    }

    // x = U^-1 c by the first wavefront, c = column N of the packed rows (or `rhs` if given):
    // lane t holds rows t and t+64; x_j is broadcast with v_readlane and subtracted from every
    // row above it (column order, j descending); returns this lane's rows in acc
    static __device__ __forceinline__ void back(const R* sU, R (&acc)[(N + 63) / 64], int lane)
    {
        // an opaque lane index (range restored): the addresses of this solve are formed here
        // rather than kept live from the kernel entry (the mixed kernel's refinement calls it
        // up to four times per iteration)
        if constexpr (NET != 14)
        {
            asm volatile("" : "+v"(lane));
            lane &= 63;
        }
        // the U entries of four columns are loaded ahead of their four dependent steps (a
        // scheduling barrier keeps each group's loads together), so the v_readlane/FMA chain does
        // not wait for LDS at every step (groups of 4 measured faster than 8 and 16: fewer live
        // registers)
        constexpr int NA = (N + 63) / 64, G = 4;
#pragma unroll
        for (int j0 = N - 1; j0 > 0; j0 -= G)
        {
            R u[G][NA];
#pragma unroll
            for (int s = 0; s < G; ++s)
#pragma unroll
                for (int a = 0; a < NA; ++a)
                {
                    const int j = j0 - s, row = lane + 64 * a;
                    // the index is selected before the load, so the address stays inside the
                    // packed factor also for lanes the condition excludes (rows at or below the
                    // diagonal, j <= 0)
                    const bool ok = j > 0 && row < j;
                    const R    x  = sU[ok ? uidx(row, j) : 0];
                    u[s][a]       = ok ? x : R(0);
                }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int s = 0; s < G; ++s)
            {
                const int j = j0 - s;
                if (j > 0)
                {
                    const R xj = bcast_lane(acc[j >> 6], j & 63);
#pragma unroll
                    for (int a = 0; a <= (j - 1) >> 6; ++a)
                    {
                        const int row = lane + 64 * a;
                        if (row < j) acc[a] = fma(-u[s][a], xj, acc[a]);
                    }
                }
            }
        }
    }
    // y = forward elimination of r with the stored multipliers (eigen.hpp:182-199 applied to a new
    // right-hand side): y_j = (r_j - sum_{k<j} L_jk y_k) / d_j, in column order; loads staged as
    // in back()
    static __device__ __forceinline__ void forward(const R* sL, const R* sRinv, R (&acc)[(N + 63) / 64], int lane)
    {
        if constexpr (NET != 14)
        {
            asm volatile("" : "+v"(lane)); // as in back()
            lane &= 63;
        }
        constexpr int NA = (N + 63) / 64, G = 4;
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += G)
        {
            R l[G][NA], ri[G];
#pragma unroll
            for (int s = 0; s < G; ++s)
            {
                const int j = j0 + s;
                ri[s]       = j < N ? sRinv[j] : R(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
                {
                    const int row = lane + 64 * a;
                    const bool ok = j < N && row > j && row < N; // index selected as in back()
                    const R    x  = sL[ok ? lrow(row) + j : 0];
                    l[s][a]       = ok ? x : R(0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int s = 0; s < G; ++s)
            {
                const int j = j0 + s;
                if (j < N)
                {
                    const R yj = bcast_lane(acc[j >> 6], j & 63) * ri[s];
                    if ((j & 63) == lane) acc[j >> 6] = yj;
#pragma unroll
                    for (int a = (j + 1) >> 6; a < NA; ++a)
                    {
                        const int row = lane + 64 * a;
                        if (row > j && row < N) acc[a] = fma(-l[s][a], yj, acc[a]);
                    }
                }
            }
        }
    }
};

// Hand-fitted rate group computed by lane tid (-1: none).
template<int NET>
__device__ __forceinline__ int special_task(int tid)
{
    // net86/net87: groups 0, 1 and 3 on wavefront 0 and groups 2 and 5 on wavefront 1 (next to
    // the fitted rates; net87's electron captures also run on wavefront 1), group 4 on wavefront 2
    // (fitted rates + species); wavefront 3 runs the species and the EOS lane.  These placements
    // measured the shortest rate phase (its slowest wavefront sets its length).
    if (NET == 86) return (tid == 32 || tid == 33) ? tid - 32 : tid == 96 ? 2 : tid == 34 ? 3 : tid == 160 ? 4 : tid == 98 ? 5 : -1;
    if (NET == 87) return (tid == 32 || tid == 33) ? tid - 32 : tid == 96 ? 2 : tid == 34 ? 3 : tid == 160 ? 4 : tid == 99 ? 5 : -1;
    return -1; // net14: the lane-parallel form on its one wavefront
}

// ---------------------------------------------------------------------------------------------
// The solve: the Cfg<NET>::NT lanes of one particle group take particles from the batch's work
// queue and integrate each over one hydro step, until the queue is empty.  sm: the group's
// Lds<NET, MIX>::total doubles of LDS; simg: the network's assembly image, staged in LDS by the
// caller; s_pf: two LDS words of the group; tid: the lane's index in the group.  WAVE: the group
// is one wavefront that shares its workgroup with other groups (wavefront barriers only).
// EosT / BatchT: the EOS and batch descriptions (EosDev / BatchDev, or views of the same fields).
// ---------------------------------------------------------------------------------------------
template<int NET, int EOSK, bool MIX, bool WAVE, class EosT, class BatchT>
__device__ __forceinline__ void solve_particles(const NetDev& nd, const EosT& ed, const BatchT& bd,
                                                const StepParams& sp, double* sm, const uint32_t* simg,
                                                unsigned* s_pf, int tid)
{
    using C  = Cfg<NET>;
    using L  = Lds<NET, MIX>;
    constexpr int D = C::D, N = C::N, NT = C::NT, WR = C::WR, TI = C::TT / C::WR, NA = (N + 63) / 64;
    constexpr int Dr = (NET == 14) ? 14 : 86; // species of the rate functor's BE/screening loops
    static_assert(!WAVE || NT == 64, "wavefront barriers: one-wavefront particle groups only");

    unsigned& s_p    = s_pf[0];
    unsigned& s_flag = s_pf[1];
    double* sBE   = sm + L::oBE;
    double* sMu   = sm + L::oMu;
    double* sEff  = sm + L::oEff;
    double* sDeff = sm + L::oDeff;
    double* sL    = sm + L::oL;
    double* sDl   = sm + L::oDl;
    double* sRate = sm + L::oRate;
    double* sDrt  = sm + L::oDrate;
    double* sVal  = sm + L::oVal;
    double* sDval = sm + L::oDval;
    double* sTr   = sm + L::oTr;
    double* sJok  = sm + L::oJok;
    double* sDYd  = sm + L::oDYdot;
    double* sDYT  = sm + L::oDYdT;
    double* sCorr = sm + L::oCorr;
    double* sEv   = sm + L::oEv;
    double* sEos  = sm + L::oEos;
    double* sJv   = sm + L::oJv;
    double* sU    = sm;
    double* sY    = sm + L::oY;
    double* sYb   = sm + L::oYb;
    double* sX    = sm + L::oX;
    double* sPan  = sm + L::oPan;
    double* sRow0 = sm + L::oRow0;

    const int lr = (tid & 63) >> 4, lc = tid & 15, wr = (tid >> 6) / WR, wc = (tid >> 6) % WR;

    const uint32_t* rx     = simg + nd.oRx;
    const uint16_t* ent16  = reinterpret_cast<const uint16_t*>(simg) + nd.oEnt16;
    const uint16_t* spp16  = reinterpret_cast<const uint16_t*>(simg) + nd.oSpPtr16;
    const uint32_t* spent  = simg + nd.oSpEnt;
    const uint16_t* nzp16  = reinterpret_cast<const uint16_t*>(simg) + nd.oNzPtr16;
    const uint32_t* nzent  = simg + nd.oNzEnt;
    const uint16_t* jord16 = reinterpret_cast<const uint16_t*>(simg) + nd.oJord16;
    const uint16_t* maperm = reinterpret_cast<const uint16_t*>(simg) + nd.oMaPerm16;
    const uint16_t* rowp16 = reinterpret_cast<const uint16_t*>(simg) + nd.oRowPtr16;
    const uint32_t* rowent = simg + nd.oRowEnt;

    const size_t q0 = bd.qrange ? (size_t)bd.qrange[0] : 0, q1 = bd.qrange ? (size_t)bd.qrange[1] : bd.n;
    for (;;)
    {
        psync<WAVE>();
        if (tid == 0) s_p = atomicAdd(bd.queue, 1u);
        psync<WAVE>();
        if (q0 + s_p >= q1) break;
        const size_t p = bd.order ? bd.order[q0 + s_p] : s_p;

        const double rho = bd.rho[p];
        double       T   = bd.temp[p];
        double       dt  = bd.dt[p];

        // parallel_nuclear_net_impl.cpp:83
        if (!(rho > k::minRho && T > k::minTemp))
        {
            if (tid == 0)
            {
                if (bd.status) bd.status[p] = NNET_STATUS_SKIPPED;
                if (bd.substeps) bd.substeps[p] = 0;
                if (bd.newton) bd.newton[p] = 0;
                if (bd.fp64_solves) bd.fp64_solves[p] = 0;
            }
            continue;
        }
        // NNET_DRHODT_CPU: parallel_nuclear_net_impl.cpp:89-92 (rho_m1 is overwritten, the
        // integration uses drho/dt = 0); NNET_DRHODT_INTEGRATE: parallel_nuclear_net.cu:79-81 (drho/dt
        // from rho_m1, which is left as it is, drives rho(t), nuclear_net.hpp:859)
        double drho_dt = 0;
        if (sp.use_drhodt == NNET_DRHODT_CPU)
        {
            if (tid == 0 && bd.rho_m1[p] != 0) bd.rho_m1[p] = (rho - bd.rho_m1[p]) / sp.previous_dt;
        }
        else if (sp.use_drhodt == NNET_DRHODT_INTEGRATE)
        {
            const double rm1 = bd.rho_m1[p];
            if (rm1 != 0) drho_dt = (rho - rm1) / sp.previous_dt;
        }

        for (int s = tid; s < D; s += NT)
            sY[s] = bd.Y[s][p];

        const double dt_tot  = sp.hydro_dt;
        double       elapsed = 0;
        double       T_buf   = __builtin_nan("");
        double       att_dt = 0, att_Tb = 0;
        long         total_iter = 0;
        uint32_t     n_sub = 0, n_newton = 0;
        bool         abandon = false; // mixed pass: FP32 refinement failed, left to the FP64 redo pass
        uint8_t      status = NNET_STATUS_OK;

        for (int i = 1;; ++i)
        {
            if (i == 1)
            {
                att_dt = dt;
                att_Tb = T_buf;
            }
            // prepareSystemSubstep (nuclear_net.hpp:858-879)
            const double rho_t   = rho - drho_dt * (dt_tot - elapsed);
            double       used_dt = dt;
            if (dt_tot - elapsed < dt) used_dt = dt_tot - elapsed;
            if (used_dt == 0)
            {
                status = NNET_STATUS_ZERO_DT;
                break;
            }
            // prepareSystemNR (nuclear_net.hpp:699-712)
            const double T_loc   = (i <= 1) ? T : T_buf;
            const double T_theta = (1 - k::theta) * T + k::theta * T_loc;
            for (int s = tid; s < D; s += NT)
            {
                const double yb = (i <= 1) ? sY[s] : sYb[s];
                sYb[s]          = (1 - k::theta) * sY[s] + k::theta * yb;
            }
            psync<WAVE>();

            // ---- EOS (one lane) and rates, phase 1 (all lanes) ----
            if constexpr (EOSK != NNET_EOS_IDEAL)
            {
                if (tid == NT - 1)
                {
                    // helmholtz.hpp:221-222: sum(Y) and sum(Y Z), each in index order; eight
                    // abundances are loaded ahead of the two interleaved addition chains
                    double abar = 0., zbar = 0;
                    int    s    = 0;
                    for (; s + 8 <= ed.zdim; s += 8)
                    {
                        double y[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            y[u] = sYb[s + u];
                        __builtin_amdgcn_sched_barrier(0); // keep the eight loads in flight together
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                        {
                            abar += y[u];
                            zbar += y[u] * ed.Z[s + u];
                        }
                    }
                    for (; s < ed.zdim; ++s)
                    {
                        abar += sYb[s];
                        zbar += sYb[s] * ed.Z[s];
                    }
                    double c0, d0;
                    helmholtz_cv_dpdt(ed.grid, ed.cells, abar, zbar, T_theta, rho_t, c0, d0);
                    sEos[0] = c0;
                    sEos[1] = d0;
                }
            }
            {
                const double Tg = T_theta;
                const int    SS = nd.nfit; // species lanes follow the fitted-rate lanes
                if (tid < nd.nfit)
                {
                    const int     f  = tid;
                    const double* fc = nd.fit + 8 * f;
                    const double  t9 = Tg / 1.e9, t913 = pow(t9, 1. / 3.), t923 = t913 * t913, t953 = pow(t9, 5. / 3.);
                    const double  t9i = 1.e0 / t9, t9i2 = t9i * t9i, t9i13 = 1.e0 / t913, t9i23 = t9i13 * t9i13,
                                 t9i43 = t9i23 * t9i23, lt9 = log(t9);
                    const double coef =
                        fc[1] * t9i + fc[2] * t9i13 + fc[3] * t913 + fc[4] * t9 + fc[5] * t953 + fc[6] * lt9;
                    const double dcoef =
                        (-fc[1] * t9i2 + (-fc[2] * t9i43 + fc[3] * t9i23 + 5 * fc[5] * t923) * (1. / 3.) + fc[4] +
                         fc[6] * t9i) *
                        1e-9;
                    const int    e   = nd.fit_e0 + f;
                    const double eff = exp(fc[0] + coef);
                    sEff[e]          = eff;
                    sDeff[e]         = eff * dcoef;
                    // reverse rate (net14.cpp:150-190, net86.cpp:524-554)
                    const double rt9  = nd.rev_div ? Tg / 1.e9 : Tg * 1e-09;
                    const double rt9i = 1. / rt9;
                    const double val1 = 11.6045 * rt9i;
                    const double val3 = val1 * rt9i * 1e-9;
                    const int    kb   = temperature_bin(Tg);
                    const double part = nd.choose[nd.ca[f] * 24 + kb] / nd.choose[nd.cb[f] * 24 + kb];
                    double       l, dl;
                    if (nd.photo[f])
                    {
                        const double val2 = 1.5 * log(rt9);
                        const double val4 = 1.5e-9 * rt9i;
                        l                 = part * exp(fc[7] + coef - val1 * nd.q[f] + val2);
                        dl                = l * (dcoef + val3 * nd.q[f] + val4);
                    }
                    else
                    {
                        l  = part * exp(fc[7] + coef - val1 * nd.q[f]);
                        dl = l * (dcoef + val3 * nd.q[f]);
                    }
                    sL[e]  = l;
                    sDl[e] = dl;
                }
                else if (tid >= SS && tid < SS + Dr)
                {
                    // corrected binding energy (net14.cpp:49-75) and screening mu/kT (:577-600)
                    const int    s          = tid - SS;
                    const double kbt        = phys::Kb * Tg;
                    const double nakbt      = phys::Na * kbt;
                    const double correction = -1.5 * nakbt;
                    const double ne         = rho_t * phys::Na / 2.;
                    {
                        const double ae  = pow((3. / 4.) / (phys::pi * ne), 1. / 3.);
                        const double gam = phys::e2 / (kbt * ae);
                        const double gmm = gam * nd.Z53[s];
                        const double fn  = gmm > 1 ? coulomb_strong(gmm) : coulomb_weak(gmm);
                        double       be  = nd.BE0[s] + correction;
                        be -= nakbt * fn;
                        sBE[s] = be;
                    }
                    {
                        const double kbt2 = Tg * phys::Kb;
                        const double ae   = pow(3. / (4. * phys::pi * ne), 1. / 3.);
                        const double gam  = phys::e2 / (kbt2 * ae);
                        const double a1 = -.898004, b1 = .96786, c1 = .220703, d1 = -.86097, e1 = 2.520058332;
                        const double gp  = gam * nd.Z53[s];
                        const double sq  = sqrt(gp);
                        const double sq2 = sqrt(sq);
                        double       mu;
                        if (gp <= 1)
                            mu = -(1. / 1.7320508075688772) * gp * sq + pow(gp, 1.9885) * .29561 / 1.9885;
                        else
                            mu = a1 * gp + 4. * (b1 * sq2 - c1 / sq2) + d1 * log(gp) - e1;
                        if (NET != 14 && s == 1) mu = 0; // neutron (net86.cpp:603)
                        sMu[s] = mu;
                    }
                }
                // hand-fitted rates.  net14: the groups' exp / pow values spread over the lanes of
                // its one wavefront, then combined in every lane; net86 / net87: one group per lane
                // on lanes of different wavefronts (measured faster there than the lane-parallel
                // form on one of their wavefronts)
                if constexpr (NET == 14)
                {
                    if ((tid >> 6) == 0)
                    {
                        // the values pass through the Jacobian array, which is rewritten only after
                        // this phase: sJv[0 .. N1) level 1, sJv[N1 .. N1 + NE2) level 2
                        static_assert(C::NNZ >= spx::N1 + spx::NE2, "hand-fitted values fit the Jacobian array");
                        const int      k = tid & 63;
                        const RatePow  p(Tg);
                        const DerivPow q(Tg);
                        {
                            const double t1 = special_level1(k, p, q);
                            if (k < spx::N1) sJv[k] = t1;
                        }
                        psync<true>(); // one wavefront
                        auto v1 = [&](int j) __attribute__((always_inline)) -> double { return sJv[j]; };
                        {
                            const double t2 = special_level2(k, p, q, v1);
                            if (k < spx::NE2) sJv[spx::N1 + k] = t2;
                        }
                        psync<true>();
                        auto v2 = [&](int j) __attribute__((always_inline)) -> double { return sJv[spx::N1 + j]; };
                        // lane 0 stores each group's rate (and reverse) with its derivative
                        auto store = [&](int g, double r, double rl, double dr, double drl) __attribute__((always_inline))
                        {
                            if (k != 0) return;
                            const int e = g == 0 ? nd.sp.r3a_e : g == 1 ? nd.sp.r24 : g == 2 ? nd.sp.r1216 : g == 3 ? nd.sp.r32
                                        : g == 4 ? nd.sp.rcag_e : nd.sp.roag_e;
                            sEff[e]  = r;
                            sDeff[e] = dr;
                            if (g == 0 || g >= 4)
                            {
                                const int l = g == 0 ? nd.sp.r3a_l : g == 4 ? nd.sp.rcag_l : nd.sp.roag_l;
                                sL[l]       = rl;
                                sDl[l]      = drl;
                            }
                        };
                        special_combine(Tg, p, q, v1, v2, store);
                    }
                }
                else
                {
                    // the t9 powers (functions of T alone) once per wavefront that holds a group,
                    // in wave-uniform code, instead of once per group on its own lane
                    static_assert(C::NT == 256, "groups on wavefronts 0-2 of four");
                    if ((tid >> 6) < 3)
                    {
                        const RatePow  p(Tg);
                        const DerivPow q(Tg);
                        const int g = special_task<NET>(tid);
                        if (g >= 0)
                        {
                            double r = 0, rl = 0, dr = 0, drl = 0;
                            if (g == 0)
                            {
                                special_triple_alpha(Tg, p, q, r, rl, dr, drl);
                                sEff[nd.sp.r3a_e] = r, sDeff[nd.sp.r3a_e] = dr, sL[nd.sp.r3a_l] = rl, sDl[nd.sp.r3a_l] = drl;
                            }
                            else if (g == 1)
                            {
                                special_c12c12(Tg, p, q, r, dr);
                                sEff[nd.sp.r24] = r, sDeff[nd.sp.r24] = dr;
                            }
                            else if (g == 2)
                            {
                                special_c12o16(Tg, p, q, r, dr);
                                sEff[nd.sp.r1216] = r, sDeff[nd.sp.r1216] = dr;
                            }
                            else if (g == 3)
                            {
                                special_o16o16(Tg, p, q, r, dr);
                                sEff[nd.sp.r32] = r, sDeff[nd.sp.r32] = dr;
                            }
                            else if (g == 4)
                            {
                                special_c12ag(Tg, p, q, r, rl, dr, drl);
                                sEff[nd.sp.rcag_e] = r, sDeff[nd.sp.rcag_e] = dr, sL[nd.sp.rcag_l] = rl,
                                sDl[nd.sp.rcag_l] = drl;
                            }
                            else
                            {
                                special_o16ag(Tg, p, q, r, rl, dr, drl);
                                sEff[nd.sp.roag_e] = r, sDeff[nd.sp.roag_e] = dr, sL[nd.sp.roag_l] = rl,
                                sDl[nd.sp.roag_l] = drl;
                            }
                        }
                    }
                }
                if (NET == 87 && (tid >> 6) == 1) // wavefront 1 (fitted rates, groups 2 and 5)
                    electron_interpolate_wave(nd.elogT, nd.elogR, nd.etab, Tg, sYb[86] * rho_t, sEv, tid & 63);
                if (NET != 14 && tid == 225) sEff[3] = sDeff[3] = sL[3] = sDl[3] = 0; // unused slots
            }
            psync<WAVE>();

            // ---- rates, phase 2: swap, screening, reaction-list order ----
            {
                const double Tg = T_theta;
                if (tid < nd.nE)
                {
                    const int i = tid;
                    double    e = sEff[i], de = sDeff[i], l = sL[i], dl = sDl[i];
                    if (nd.swap[i])
                    {
                        double t = e;
                        e        = l;
                        l        = t;
                        t        = de;
                        de       = dl;
                        dl       = t;
                    }
                    const int* sc = nd.scr + 4 * i;
                    double     dm = sMu[sc[0]] + sMu[sc[1]] - sMu[sc[2]];
                    // the optional fourth species (-1: none) is loaded through a selected index,
                    // so a load issued ahead of the test still addresses sMu
                    const double mu3 = sMu[sc[3] >= 0 ? sc[3] : 0];
                    if (sc[3] >= 0) dm += mu3;
                    const double EF = exp(dm);
                    e               = e * EF;
                    de              = de * EF - 2. * e * dm / Tg;
                    if (nd.scr_l[i])
                    {
                        l  = l * EF;
                        dl = dl * EF - 2. * l * dm / Tg;
                    }
                    if (nd.push_e[i] >= 0) sRate[nd.push_e[i]] = e, sDrt[nd.push_e[i]] = de;
                    if (nd.push_l[i] >= 0) sRate[nd.push_l[i]] = l, sDrt[nd.push_l[i]] = dl;
                }
                if (NET == 87 && tid == NT - 1)
                {
                    // net87.hpp:153-174
                    const double kbt        = phys::Kb * Tg;
                    const double nakbt      = phys::Na * kbt;
                    const double correction = -1.5 * nakbt;
                    sBE[86]                 = nd.BEe + correction;
                    sBE[0] += -sEv[3];
                    sBE[1] += -sEv[9];
                    const double rhoElec = sYb[86] * rho_t;
                    sRate[309]           = rhoElec == 0 ? 0 : sEv[0] / rhoElec;
                    sDrt[309]            = rhoElec == 0 ? 0 : sEv[1] / rhoElec;
                    sRate[310]           = rhoElec == 0 ? 0 : sEv[6] / rhoElec;
                    sDrt[310]            = rhoElec == 0 ? 0 : sEv[7] / rhoElec;
                }
            }
            psync<WAVE>();

            // ---- per-reaction mass-action factors (nuclear_net.hpp:355-371, 404-426) ----
            // slot s computes reaction maperm[s]: the wavefronts that run a second pass take the
            // cheapest reactions
            for (int s = tid; s < nd.nR; s += NT)
            {
                const int    r = maperm[s];
                double       v = sRate[r], dv = sDrt[r], jr = sRate[r];
                const int    q0 = rx[r] & 0xffff, q1 = rx[r] >> 16;
                for (int q = q0; q < q1; ++q)
                {
                    const int    c = ent16[q] >> 8;
                    const double x = sYb[ent16[q] & 0xff] * rho_t;
                    if (c != 1)
                    {
                        const double fct = factorial_count(c);
                        v                = v / fct;
                        dv               = dv / fct;
                        jr               = jr / fct;
                    }
                    v  = v * pow_count(x, c);
                    dv = dv * pow_count(x, c);
                    jr = jr * pow_count(x, c - 1);
                }
                sVal[r]       = v / rho_t;
                sDval[r]      = dv / rho_t;
                const bool ok = jr > k::epsilonSystem;
                sJok[r]       = ok ? 1. : 0.;
                if (ok)
                    for (int qa = q0; qa < q1; ++qa)
                    {
                        const int b   = ent16[qa] & 0xff;
                        double    tr_ = jr;
                        for (int o = q0; o < q1; ++o)
                            if ((ent16[o] & 0xff) != b) tr_ *= sYb[ent16[o] & 0xff] * rho_t;
                        sTr[r * MAXQ + (qa - q0)] = tr_;
                    }
            }
            psync<WAVE>();

            // ---- Jacobian non-zeros (firstOrderDYFromReactions) and species sums
            //      (derivativesFromReactions), each summed in the reference's order ----
            // non-zeros are handed out in descending order of their list lengths, so the lanes of
            // one pass over the workgroup have lists of similar length (lane-to-slot map: jac_slot)
            for (int pass = 0;; ++pass)
            {
                const int k = jac_slot(tid, pass, NT, nd.nnz);
                if (k >= nd.nnz) break;
                const int e = jord16[k];
                // J = sum over the entry's contributions in list order of count * mass-action
                // derivative, skipping reactions whose derivative factor is <= 1e-40; eight
                // contributions are loaded ahead of their (sequential) additions
                double    J  = 0.;
                int       kk = nzp16[e];
                const int k1 = nzp16[e + 1];
                for (; kk + 8 <= k1; kk += 8)
                {
                    // stage 1: the eight entry words; stage 2: their factors; then the additions
                    // (scheduling barriers keep each stage's loads in flight together)
                    uint32_t w[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        w[u] = nzent[kk + u];
                    __builtin_amdgcn_sched_barrier(0);
                    double tr[8], okv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const int r = w[u] & 0xfff;
                        okv[u]      = sJok[r];
                        tr[u]       = sTr[r * MAXQ + ((w[u] >> 12) & 0xf)];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const double t = tr[u] * (double)(int8_t)(w[u] >> 16);
                        J              = okv[u] != 0. ? J + t : J;
                    }
                }
                for (; kk < k1; ++kk)
                {
                    const uint32_t w  = nzent[kk];
                    const int      r  = w & 0xfff;
                    const double   t  = sTr[r * MAXQ + ((w >> 12) & 0xf)] * (double)(int8_t)(w >> 16);
                    J                 = sJok[r] != 0. ? J + t : J;
                }
                sJv[e] = J;
            }
            // dY/dt and d(dY/dt)/dT per species: signed-count sums in list order.  Nothing reads
            // them before the matrix fill, so with several wavefronts per particle each sum takes
            // its first KPRE entries here and the rest after the barrier, beside the RHS
            // correction of the other lanes (the longest species lists -- 140 entries for two
            // net86 species -- would otherwise hold every wavefront at the barrier); one
            // wavefront has nothing to overlap.  The additions keep their order across the split.
            constexpr bool SPLIT = C::NT > 64;
            constexpr int  KPRE  = SPLIT ? 72 : 1 << 20; // 72 measured faster than 96 on net86
            static_assert(2 * D <= NT, "species lanes [NT - D, NT) and RHS lanes [0, D) are disjoint");
            const int     sps  = tid - (NT - D);
            double        spa = 0., spb = 0.;
            int           spe = 0, spe1 = 0;
            auto species_sum = [&](int e1) __attribute__((always_inline))
            {
                int e = spe;
                for (; e + 8 <= e1; e += 8)
                {
                    uint32_t w[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        w[u] = spent[e + u];
                    __builtin_amdgcn_sched_barrier(0);
                    double x[8], y[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const int r = w[u] & 0xffff;
                        x[u]        = sVal[r];
                        y[u]        = sDval[r];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        const double c = (double)(int8_t)(w[u] >> 16);
                        spa            = spa + x[u] * c;
                        spb            = spb + y[u] * c;
                    }
                }
                for (; e < e1; ++e)
                {
                    const uint32_t w = spent[e];
                    const int      r = w & 0xffff;
                    const double   c = (double)(int8_t)(w >> 16);
                    spa              = spa + sVal[r] * c;
                    spb              = spb + sDval[r] * c;
                }
                spe = e;
            };
            // Yg - Y per species for the RHS correction, into the screening array (dead after the
            // rate phase): the difference the reference forms inside that sum, formed once here
            double* sDYg = sMu;
            if (sps >= 0)
            {
                sDYg[sps] = sYb[sps] - sY[sps];
                spe  = spp16[sps];
                spe1 = spp16[sps + 1];
                species_sum(spe1 - spe > KPRE ? spe + KPRE : spe1);
                if constexpr (!SPLIT)
                {
                    sDYd[sps] = spa;
                    sDYT[sps] = spb;
                }
            }
            // this lane's register-slot mask and first non-zero id (for the matrix fill)
            const uint4 lmap = nd.lanemap[tid];
            psync<WAVE>();
            // RHS correction sum_j J[a][j] (Yg[j] - Y[j]) over the row's non-zeros, j ascending
            if (tid < D)
            {
                double    corr = 0;
                int       e    = rowp16[tid];
                const int e1   = rowp16[tid + 1];
                for (; e + 8 <= e1; e += 8)
                {
                    uint32_t w[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        w[u] = rowent[e + u];
                    __builtin_amdgcn_sched_barrier(0);
                    double jv[8], dy[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                    {
                        jv[u] = sJv[w[u] & 0xffff];
                        dy[u] = sDYg[w[u] >> 16];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        corr = corr + jv[u] * dy[u];
                }
                for (; e < e1; ++e)
                {
                    const uint32_t w = rowent[e];
                    corr             = corr + sJv[w & 0xffff] * sDYg[w >> 16];
                }
                sCorr[tid] = corr;
            }
            else if (SPLIT && sps >= 0)
            {
                species_sum(spe1);
                sDYd[sps] = spa;
                sDYT[sps] = spb;
            }
            psync<WAVE>();

            // ---- Newton matrix into registers (nuclear_net.hpp:551-608) ----
            double cvv, dpdTv;
            if constexpr (EOSK == NNET_EOS_IDEAL)
            {
                const double dmy = 8.317e7 / ed.mu;
                cvv              = 1.5 * dmy;
                dpdTv            = rho_t * dmy;
            }
            else
            {
                cvv   = sEos[0];
                dpdTv = sEos[1];
            }
            const double mthdt   = -k::theta * used_dt;
            const double thdt    = k::theta * used_dt;
            const double drho    = drho_dt * used_dt;
            const double value_1 = dpdTv * drho / (rho_t * rho_t);
            // matrix entry (ri, ci) from the assembly arrays, as prepareSystemFromGuess builds it;
            // J is the Jacobian entry (0 where the pattern has none)
            auto entry = [&](int ri, int ci, double J) __attribute__((always_inline)) -> double
            {
                if (ri == 0)
                {
                    if (ci == 0) return 1 - k::theta * value_1 / cvv;
                    if (ci < N) return -sBE[ci - 1] / cvv;
                    return T * value_1 / cvv;
                }
                const int s = ri - 1;
                if (ci == 0) return sDYT[s] * mthdt;
                if (ci < N) return (ci == ri) ? 1. - thdt * J : mthdt * J;
                double rhs = sDYd[s] * used_dt;
                rhs += mthdt * sDYT[s] * (T_theta - T);
                rhs += mthdt * sCorr[s];
                return rhs;
            };
            // the lane's register slots (tile placement of TileLU) filled from the Jacobian
            // non-zeros, which are numbered lane-major in slot order
            // Every slot first takes its Jacobian-block value (rows and columns 1 .. N-1, zero
            // elsewhere) without branches; then the few slots of column 0 (dY/dT), column N (the
            // right-hand side) and row 0 (temperature) are overwritten where they sit: column 0 in
            // register lane column 0 of tile column 0, column N in lane column N & 15 of tile column
            // N >> 4, row 0 in register 0 of lane row 0 of tile row 0.  Each slot gets the value
            // entry() gives it.
            auto fill = [&](auto& Mt) __attribute__((always_inline))
            {
                const uint64_t mask = (uint64_t)lmap.x | ((uint64_t)lmap.y << 32);
                const int      base = (int)lmap.z; // id of this lane's first non-zero
#pragma unroll
                for (int ti = 0; ti < TI; ++ti)
#pragma unroll
                    for (int tj = 0; tj < TI; ++tj)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            const int  slot = (ti * TI + tj) * 4 + v;
                            const int  ri = tile_row(wr + WR * ti, lr, v), ci = tile_col(wc + WR * tj, lc);
                            const bool nz = (mask >> slot) & 1;
                            // the lane's non-zeros are numbered in slot order
                            const int    nzi = base + (int)__popcll(mask & ((1ull << slot) - 1ull));
                            const double Jl  = sJv[nz ? nzi : 0];
                            const double J   = nz ? Jl : 0.;
                            const double x   = (ci == ri) ? 1. - thdt * J : mthdt * J;
                            Mt[ti][tj][v]    = (ri >= 1 && ri < N && ci >= 1 && ci < N) ? x : 0.;
                        }
                // column 0: dY/dT
                if (wc == 0 && lc == 0)
#pragma unroll
                    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            const int ri = tile_row(wr + WR * ti, lr, v);
                            if (ri >= 1 && ri < N) Mt[ti][0][v] = entry(ri, 0, 0.);
                        }
                // column N: the right-hand side
                constexpr int JN = N >> 4;
                if (wc == JN % WR && lc == (N & 15))
#pragma unroll
                    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                        {
                            const int ri = tile_row(wr + WR * ti, lr, v);
                            if (ri >= 1 && ri < N) Mt[ti][JN / WR][v] = entry(ri, N, 0.);
                        }
                // row 0 (temperature)
                if (wr == 0 && lr == 0)
#pragma unroll
                    for (int tj = 0; tj < TI; ++tj)
                    {
                        const int ci = tile_col(wc + WR * tj, lc);
                        if (ci <= N)
                        {
                            const double x = entry(0, ci, 0.);
                            Mt[0][tj][0]   = x;
                            if constexpr (MIX) sRow0[ci] = x;
                        }
                    }
            };
            auto solve_fp64 = [&]() __attribute__((always_inline))
            {
                dx4 Mt[TI][TI];
                fill(Mt);
                // the packed rows the LU writes alias the arrays the fill reads: the barrier after
                // the LU's first panel publish (which writes only the panel buffers) orders every
                // lane's fill before the first packed-row store
                // ---- LU (eigen.hpp:179-206), four pivots per block ----
                TileLU<NET, double, false, WAVE>::factor(Mt, sPan, sU, nullptr, nullptr, wr, wc, lr, lc, tid);
                // ---- back substitution (eigen.hpp:208-221) ----
                if (tid < 64)
                {
                    double acc[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a)
                    {
                        const int row = tid + 64 * a;
                        const double x = sU[row < N ? TileLU<NET, double, false, WAVE>::uidx(row, N) : 0];
                        acc[a]         = row < N ? x : 0.;
                    }
                    TileLU<NET, double, false, WAVE>::back(sU, acc, tid);
#pragma unroll
                    for (int a = 0; a < NA; ++a)
                        if (tid + 64 * a < N) sX[tid + 64 * a] = acc[a];
                }
                psync<WAVE>();
            };

            // the FP64 solve has a single (inlined) call site, so that every LDS access of the
            // kernel stays a direct LDS instruction; the mixed kernel does not contain it
            if constexpr (!MIX)
                solve_fp64();
            else
            {
                // ---- mixed precision: FP32 tile LU + FP64 iterative refinement; a particle
                //      with an iteration whose refinement does not converge is abandoned here
                //      and integrated again, from the start of the step, by the FP64 pass ----
                using LU32 = TileLU<NET, float, true, WAVE>;
                float* sU32   = reinterpret_cast<float*>(sm);
                float* sL32   = sU32 + L::NU;
                float* sRi32  = sL32 + L::NL;
                float* sPan32 = reinterpret_cast<float*>(sPan);
                {
                    fx4 Mt[TI][TI];
                    fill(Mt);
                    // FP32 factors alias the head arrays only; the first is stored after the LU's
                    // first panel barrier
                    LU32::factor(Mt, sPan32, sU32, sL32, sRi32, wr, wc, lr, lc, tid);
                }
                if (tid < 64)
                {
                    float acc[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a)
                    {
                        const int row = tid + 64 * a;
                        const float x = sU32[row < N ? LU32::uidx(row, N) : 0];
                        acc[a]        = row < N ? x : 0.f;
                    }
                    LU32::back(sU32, acc, tid);
#pragma unroll
                    for (int a = 0; a < NA; ++a)
                        if (tid + 64 * a < N) sX[tid + 64 * a] = (double)acc[a];
                }
                psync<WAVE>();
                double* sRes = sPan; // the panel buffers are free after the LU
                bool    conv = false;
                for (int sweep = 0; sweep < k::mixed_sweeps(NET) && !conv; ++sweep)
                {
                    // FP64 residual r = b - M x over the sparse pattern (row 0 is dense)
                    for (int ri = tid; ri < N; ri += NT)
                    {
                        double r;
                        if (ri == 0)
                        {
                            r = sRow0[N];
                            for (int c = 0; c < N; ++c)
                                r = fma(-sRow0[c], sX[c], r);
                        }
                        else
                        {
                            const int s = ri - 1;
                            r           = entry(ri, N, 0.);
                            r           = fma(-entry(ri, 0, 0.), sX[0], r);
                            bool diag   = false;
                            for (int e = rowp16[s]; e < rowp16[s + 1]; ++e)
                            {
                                const uint32_t w = rowent[e];
                                const int      c = (int)(w >> 16) + 1;
                                diag             = diag || c == ri;
                                r                = fma(-entry(ri, c, sJv[w & 0xffff]), sX[c], r);
                            }
                            if (!diag) r -= sX[ri];
                        }
                        sRes[ri] = r;
                    }
                    psync<WAVE>();
                    // correction d = (LU)^-1 r in FP32; converged when every |d_i| is below
                    // mixedTol of its magnitude scale (|x_0| + T for the temperature, |x_i| + Y_i
                    // for the abundances)
                    if (tid < 64)
                    {
                        float acc[NA];
#pragma unroll
                        for (int a = 0; a < NA; ++a)
                        {
                            const int row = tid + 64 * a;
                            acc[a]        = row < N ? (float)sRes[row] : 0.f;
                        }
                        LU32::forward(sL32, sRi32, acc, tid);
                        LU32::back(sU32, acc, tid);
                        bool ok = true;
#pragma unroll
                        for (int a = 0; a < NA; ++a)
                        {
                            const int row = tid + 64 * a;
                            if (row < N)
                            {
                                const double d     = (double)acc[a];
                                const double xn    = sX[row] + d;
                                const double scale = fabs(xn) + (row == 0 ? T : fabs(sY[row - 1])) + 1e-30;
                                ok                 = ok && fabs(d) <= k::mixedTol * scale;
                                sX[row]            = xn;
                            }
                        }
                        const bool all = __all(ok);
                        if (tid == 0) s_flag = all ? 1u : 0u;
                    }
                    psync<WAVE>();
                    conv = s_flag != 0;
                }
                abandon = !conv;
            }
            if constexpr (MIX)
                if (abandon) break;
            ++n_newton;
            ++total_iter;

            // ---- finalizeSystemSubstep / finalizeSystemNR (nuclear_net.hpp:724-782, 887-922) ----
            double u = dt;
            if (dt_tot - elapsed < dt) u = dt_tot - elapsed;
            double       timestep = 0;
            bool         conv     = false;
            const double last_T   = T_buf;
            const double T_new    = T + sX[0];
            int          nanflag  = 0;
            double       yb_new   = 0.;
            if (tid < D)
            {
                yb_new  = sY[tid] + sX[tid + 1];
                nanflag = isnan(yb_new) ? 1 : 0;
            }
            const bool has_nan = psync_or<WAVE>(nanflag) || isnan(T_new);
            if (has_nan || T_new < 0)
            {
                u *= k::nanDtStep;
                i     = 0;
                T_buf = T_new;
                if (tid < D) sYb[tid] = yb_new;
            }
            else
            {
                const double dT_T = fabs((T_new - T) / T_new);
                if (dT_T > k::NR_dTOverTempTarget * k::NR_dTOverTempTol)
                {
                    u *= k::NR_dTOverTempTol / dT_T;
                    i     = 0;
                    T_buf = T;
                    if (tid < D) sYb[tid] = sY[tid];
                }
                else
                {
                    if (tid < D) sYb[tid] = (yb_new <= k::epsilonVector) ? 0. : yb_new;
                    T_buf                   = T_new;
                    const double correction = fabs((T_buf - last_T) / T_buf);
                    if (correction < k::NR_itTol || i >= k::NR_maxIt)
                    {
                        timestep = u;
                        u        = (dT_T == 0 ? k::maxDtStep : k::NR_dTOverTempTarget / dT_T) * timestep;
                        u        = dmin(u, timestep * (double)k::maxDtStep);
                        u        = dmax(u, timestep * (double)k::minDtStep);
                        u        = dmin(u, (double)k::NR_maxDt);
                        conv     = true;
                    }
                }
            }
            if (dt_tot - elapsed < dt)
            {
                if (u < dt_tot - elapsed) dt = u;
            }
            else
                dt = u;
            if (conv)
            {
                if (tid < D) sY[tid] = sYb[tid];
                T = T_buf;
                i = 0;
                elapsed += timestep;
                ++n_sub;
                if ((dt_tot - elapsed) / dt_tot < k::substep_dtTol) break;
            }
            else if (i == 0)
            {
                const bool same_dt = __double_as_longlong(dt) == __double_as_longlong(att_dt);
                const bool same_T  = __double_as_longlong(T_buf) == __double_as_longlong(att_Tb);
                if (same_dt && same_T)
                {
                    status = NNET_STATUS_STALLED;
                    break;
                }
            }
            if (sp.max_iter > 0 && total_iter >= sp.max_iter)
            {
                status = NNET_STATUS_MAX_ITER;
                break;
            }
            psync<WAVE>();
        }
        psync<WAVE>();
        if constexpr (MIX)
            if (abandon)
            {
                if (tid == 0) bd.redo[atomicAdd(bd.redo_count, 1u)] = (uint32_t)p;
                continue;
            }
        for (int s = tid; s < D; s += NT)
            bd.Y[s][p] = sY[s];
        if (tid == 0)
        {
            bd.temp[p] = T;
            bd.dt[p]   = dt;
            if (bd.status) bd.status[p] = status;
            if (bd.substeps) bd.substeps[p] = n_sub;
            if (bd.newton) bd.newton[p] = n_newton;
            if (bd.fp64_solves) bd.fp64_solves[p] = bd.fp64_pass ? n_newton : 0u;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// The solve kernel of one batch: persistent workgroups of PPW particle groups take particles
// from the batch's work queue.  PPW > 1 (one-wavefront networks only): the groups of a workgroup
// share the staged assembly image and otherwise run independently (wavefront barriers).
// ---------------------------------------------------------------------------------------------
template<int NET, int EOSK, bool MIX, int PPW>
__global__ void __launch_bounds__(Cfg<NET>::NT * PPW, Cfg<NET>::WPE)
    nnet_solve_kernel(NetDev nd, EosDev ed, BatchDev bd, StepParams sp)
{
    using C = Cfg<NET>;
    using L = Lds<NET, MIX>;
    static_assert(PPW == 1 || C::NT == 64, "several particles per workgroup: one-wavefront networks only");
    __shared__ __align__(16) double sm[PPW * L::total];
    __shared__ uint32_t simg[C::IMG];
    __shared__ unsigned s_pf[2 * PPW];
    const int group = PPW == 1 ? 0 : (int)threadIdx.x / C::NT;
    const int tid   = PPW == 1 ? (int)threadIdx.x : (int)threadIdx.x % C::NT;
    // stage the network's assembly image once per workgroup
    for (int w = threadIdx.x; w < nd.img_words; w += C::NT * PPW)
        simg[w] = nd.img[w];
    __syncthreads();
    solve_particles<NET, EOSK, MIX, (PPW > 1)>(nd, ed, bd, sp, sm + group * L::total, simg, s_pf + 2 * group, tid);
}

// ---------------------------------------------------------------------------------------------
// One launch for the batches of several networks with one EOS and one precision.  Each
// workgroup drains the batches' work queues phase by phase (the host orders the phases by cost:
// net87 and net86 before net14), staging each phase's assembly image; net86 / net87 particles
// take the whole workgroup, net14 particles one wavefront each (four per workgroup).  The
// charges and abundance pointers of all phases share two arrays of FUSE_SPECIES entries, which
// keeps the argument block (passed by value, so a captured HIP graph holds its own copy) under
// 4 KiB.
// ---------------------------------------------------------------------------------------------
constexpr int FUSE_PHASES = 4, FUSE_SPECIES = 192, FUSE_NT = 256;
struct PhaseDev
{
    const NetDev* nd;   // device copy of the network's NetDev
    int           kind; // 14, 86 or 87
    int           zoff; // first charge of the phase in FusedDev::Z
    int           yoff; // first abundance pointer of the phase in FusedDev::Y
    EosCore       eos;
    BatchCore     b;
};
struct FusedDev
{
    int      nphase;
    PhaseDev ph[FUSE_PHASES];
    double   Z[FUSE_SPECIES];
    double*  Y[FUSE_SPECIES];
};
static_assert(sizeof(FusedDev) + sizeof(StepParams) <= 4096, "fused kernel arguments within 4 KiB");

template<bool MIX>
struct FusedLds
{
    static constexpr int m2(int a, int b) { return a > b ? a : b; }
    static constexpr int total =
        m2(m2(Lds<86, MIX>::total, Lds<87, MIX>::total), (FUSE_NT / 64) * Lds<14, MIX>::total);
    static constexpr int img = m2(m2(Cfg<86>::IMG, Cfg<87>::IMG), Cfg<14>::IMG);
};
static_assert(Cfg<86>::NT == FUSE_NT && Cfg<87>::NT == FUSE_NT && Cfg<14>::NT == 64, "fused workgroup shape");

template<int EOSK, bool MIX>
__global__ void __launch_bounds__(FUSE_NT, 2) nnet_solve_fused(FusedDev fd, StepParams sp)
{
    using F = FusedLds<MIX>;
    __shared__ __align__(16) double sm[F::total];
    __shared__ uint32_t simg[F::img];
    __shared__ unsigned s_pf[2 * (FUSE_NT / 64)];
    for (int ph = 0; ph < fd.nphase; ++ph)
    {
        const PhaseDev& P  = fd.ph[ph];
        const NetDev&   nd = *P.nd;
        EosView         ev;
        static_cast<EosCore&>(ev) = P.eos;
        ev.Z                      = fd.Z + P.zoff;
        BatchView bv;
        static_cast<BatchCore&>(bv) = P.b;
        bv.Y                        = fd.Y + P.yoff;
        __syncthreads(); // every group is done with the previous phase's image and LDS
        for (int w = threadIdx.x; w < nd.img_words; w += FUSE_NT)
            simg[w] = nd.img[w];
        __syncthreads();
        // the lane index enters each network's code through an empty asm statement, so the
        // compiler forms each network's lane-dependent values inside its own branch instead of
        // hoisting common ones above the branches, where they would stay live through all three
        // (fused Helmholtz FP64 kernel: 72 -> 25 spilled VGPRs)
        int t = threadIdx.x;
        if (P.kind == 87)
        {
            asm volatile("" : "+v"(t));
            solve_particles<87, EOSK, MIX, false>(nd, ev, bv, sp, sm, simg, s_pf, t);
        }
        else if (P.kind == 86)
        {
            asm volatile("" : "+v"(t));
            solve_particles<86, EOSK, MIX, false>(nd, ev, bv, sp, sm, simg, s_pf, t);
        }
        else
        {
            asm volatile("" : "+v"(t));
            const int g = t >> 6;
            solve_particles<14, EOSK, MIX, true>(nd, ev, bv, sp, sm + g * Lds<14, MIX>::total, simg, s_pf + 2 * g,
                                                 t & 63);
        }
    }
}

// The solve kernels are compiled in translation units of their own, one per network (this file
// built with NNET_SOLVE_NET=<network>, which keeps only the code above), so the library's
// objects build in parallel; the host part below launches them through these declarations.
#define NNET_SOLVE_INSTANCE(SPEC, NET, PPW)                                                                           \
    SPEC template __global__ void nnet_solve_kernel<NET, NNET_EOS_IDEAL, false, PPW>(NetDev, EosDev, BatchDev,         \
                                                                                     StepParams);                     \
    SPEC template __global__ void nnet_solve_kernel<NET, NNET_EOS_IDEAL, true, PPW>(NetDev, EosDev, BatchDev,          \
                                                                                    StepParams);                      \
    SPEC template __global__ void nnet_solve_kernel<NET, NNET_EOS_HELMHOLTZ, false, PPW>(NetDev, EosDev, BatchDev,     \
                                                                                         StepParams);                 \
    SPEC template __global__ void nnet_solve_kernel<NET, NNET_EOS_HELMHOLTZ, true, PPW>(NetDev, EosDev, BatchDev,      \
                                                                                        StepParams);
#define NNET_SOLVE_INSTANCES(SPEC, NET)                                                                               \
    NNET_SOLVE_INSTANCE(SPEC, NET, 1)                                                                                 \
    NNET_SOLVE_INSTANCE(SPEC, NET, 2)                                                                                 \
    NNET_SOLVE_INSTANCE(SPEC, NET, 4)
// NNET_SOLVE_NET = 1000 + 2 eos + mixed: the fused kernel of that EOS and precision
#define NNET_FUSED_INSTANCE(SPEC, EOSK, MIX) SPEC template __global__ void nnet_solve_fused<EOSK, MIX>(FusedDev, StepParams);
#ifdef NNET_SOLVE_NET
#if NNET_SOLVE_NET == 14
NNET_SOLVE_INSTANCES(, 14)
#elif NNET_SOLVE_NET >= 1000
NNET_FUSED_INSTANCE(, (NNET_SOLVE_NET - 1000) / 2, (NNET_SOLVE_NET - 1000) % 2 == 1)
#else
NNET_SOLVE_INSTANCE(, NNET_SOLVE_NET, 1)
#endif
} // namespace nnet_amd
#else
NNET_SOLVE_INSTANCES(extern, 14)
NNET_SOLVE_INSTANCE(extern, 86, 1)
NNET_SOLVE_INSTANCE(extern, 87, 1)
NNET_FUSED_INSTANCE(extern, NNET_EOS_IDEAL, false)
NNET_FUSED_INSTANCE(extern, NNET_EOS_IDEAL, true)
NNET_FUSED_INSTANCE(extern, NNET_EOS_HELMHOLTZ, false)
NNET_FUSED_INSTANCE(extern, NNET_EOS_HELMHOLTZ, true)

// ---------------------------------------------------------------------------------------------
// Post-step EOS refresh (computeHelmholtzImpl, parallel_nuclear_net_impl.cpp:116-141): one lane
// per particle.  The abundances are read species by species, each load coalesced across the
// wavefront, sixteen at a time ahead of the two in-order sums (sum(Y), then sum(Y Z) as
// eigen::dot); each lane then gathers its 288 + 128-byte cell blocks of the table.  Groups of 16
// measured 2% faster than 8 and 9% faster than 32 (the kernel alternates load and EOS phases at
// 3 waves per SIMD; its HBM traffic equals the algorithmic bytes).
// ---------------------------------------------------------------------------------------------
constexpr int HELM_EOS_MAX_DIM = 128;
struct HelmEosArgs
{
    size_t        n;
    int           dim;
    const double* temp;
    const double* rho;
    double*       u;
    double*       cv;
    double*       p;
    double*       c;
    double*       dpdT;
    const double* Y[HELM_EOS_MAX_DIM];
    double        Z[HELM_EOS_MAX_DIM];
};

__global__ void __launch_bounds__(256) nnet_helm_eos_kernel(HelmEosArgs a, const HelmGrid* __restrict__ g,
                                                            const double* __restrict__ cells,
                                                            const double* __restrict__ cells3)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double abar = 0., zbar = 0.;
    int    s    = 0;
    for (; s + 16 <= a.dim; s += 16)
    {
        double y[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            y[u] = a.Y[s + u][i];
#pragma unroll
        for (int u = 0; u < 16; ++u)
        {
            abar += y[u];
            zbar += y[u] * a.Z[s + u];
        }
    }
    for (; s < a.dim; ++s)
    {
        const double y = a.Y[s][i];
        abar += y;
        zbar += y * a.Z[s];
    }
    HelmOut o;
    helmholtz_eval<true>(g, cells, cells3, abar, zbar, a.temp[i], a.rho[i], o);
    if (a.u) a.u[i] = o.u;
    if (a.cv) a.cv[i] = o.cv;
    if (a.p) a.p[i] = o.p;
    if (a.c) a.c[i] = o.c;
    if (a.dpdT) a.dpdT[i] = o.dpdT;
}

__global__ void nnet_queue_reset(unsigned int* q, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = 0u;
}

// ---------------------------------------------------------------------------------------------
// Cost-ordered dispatch: particles are handed to the persistent workgroups in descending buckets
// of an expected cost (the caller's hint, e.g. the Newton iterations of the previous call), so the
// long integrations start first and do not form a tail.  Bucket = bit length of the hint.
// ---------------------------------------------------------------------------------------------
constexpr int NBUCKET = 32;
__device__ __forceinline__ int cost_bucket(uint32_t c)
{
    const int b = c == 0 ? 0 : 32 - __clz(c);
    return b < NBUCKET ? b : NBUCKET - 1;
}

// Dispatch key: the cost bucket, plus NBUCKET for a particle routed to the FP64 pass of a mixed
// batch (nonzero fp64 hint).  Keys are laid out in descending order, so the routed particles form
// the prefix [0, nH) of the order and the others [nH, n), each longest first.
constexpr int NKEY = 2 * NBUCKET, BIN_STRIDE = 2 * NKEY + 4; // counts, offsets, range {0, nH, n}
__device__ __forceinline__ int order_key(const uint32_t* cost, const uint32_t* hint, size_t i)
{
    return (hint && hint[i] != 0u ? NBUCKET : 0) + (cost ? cost_bucket(cost[i]) : 0);
}

// bins[0 .. NKEY): counts; zeroed by nnet_queue_reset before the launch
__global__ void __launch_bounds__(256) nnet_order_count(const uint32_t* __restrict__ cost,
                                                        const uint32_t* __restrict__ hint, size_t n, unsigned* bins)
{
    __shared__ unsigned h[NKEY];
    if (threadIdx.x < NKEY) h[threadIdx.x] = 0u;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        atomicAdd(&h[order_key(cost, hint, i)], 1u);
    __syncthreads();
    if (threadIdx.x < NKEY && h[threadIdx.x]) atomicAdd(&bins[threadIdx.x], h[threadIdx.x]);
}

// bins[NKEY + k] = first position of key k (largest key first); bins[2 NKEY ..] = {0, nH, n}
__global__ void nnet_order_offsets(unsigned* bins)
{
    if (threadIdx.x == 0)
    {
        unsigned acc = 0;
        for (int k = NKEY - 1; k >= 0; --k)
        {
            if (k == NBUCKET - 1) bins[2 * NKEY + 1] = acc;
            bins[NKEY + k] = acc;
            acc += bins[k];
        }
        bins[2 * NKEY]     = 0u;
        bins[2 * NKEY + 2] = acc;
    }
}

// each 256-particle tile reserves its key ranges with one atomic per key
__global__ void __launch_bounds__(256) nnet_order_scatter(const uint32_t* __restrict__ cost,
                                                          const uint32_t* __restrict__ hint, size_t n, unsigned* bins,
                                                          uint32_t* __restrict__ order)
{
    __shared__ unsigned h[NKEY], base[NKEY];
    for (size_t t0 = (size_t)blockIdx.x * blockDim.x; t0 < n; t0 += (size_t)gridDim.x * blockDim.x)
    {
        if (threadIdx.x < NKEY) h[threadIdx.x] = 0u;
        __syncthreads();
        const size_t   i = t0 + threadIdx.x;
        const int      k = i < n ? order_key(cost, hint, i) : 0;
        const unsigned r = i < n ? atomicAdd(&h[k], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x < NKEY && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&bins[NKEY + threadIdx.x], h[threadIdx.x]);
        __syncthreads();
        if (i < n) order[base[k] + r] = (uint32_t)i;
        __syncthreads();
    }
}

} // namespace nnet_amd

// =============================================================================================
// Host side
// =============================================================================================
using namespace nnet_amd;

// Launch parameters of the solve kernels per network, EOS and precision: the values of the
// autotuner's table (sph-exa-nnet_amd/autotune.py; nnet_ctx_load_tuning) or of
// nnet_ctx_set_particles_per_workgroup.
struct Tuning
{
    // particle groups per workgroup of net14, [eos][mixed]; net86 / net87 always run one
    int ppw14[2][2] = {{1, 1}, {1, 1}};
    int ppw(int net, int eos, bool mixed) const { return net == 14 ? ppw14[eos ? 1 : 0][mixed ? 1 : 0] : 1; }
    // calls with batches of several networks: one fused launch (1) or one launch per batch on side
    // streams (0)
    int fused = 1;
};

struct nnet_ctx
{
    int         device = 0;
    std::string err;
    long        max_iter = 10000;
    Tuning      tuning;
    // Helmholtz
    HelmGrid* d_grid  = nullptr;
    double*   d_cells  = nullptr;
    double*   d_cells3 = nullptr;
    bool      helm_ok  = false;
    // NCTR counters per batch of a call: work queues of the first pass, of the routed FP64 pass
    // and of the redo pass; then the redo range {0, count}
    static constexpr int NCTR = 5;
    unsigned int* d_queue = nullptr;
    int           n_queue = 0; // batches the counters cover
    // mixed precision: the particles handed from the mixed pass to the FP64 redo pass
    uint32_t* d_redo = nullptr;
    size_t    n_redo = 0;
    int           n_cu    = 0;
    // ordered dispatch: per batch BIN_STRIDE key counts, offsets and ranges, and the particle order
    unsigned int* d_bins   = nullptr;
    uint32_t*     d_order  = nullptr;
    size_t        n_order  = 0;
    int           n_bins   = 0;
    // batches of one call run concurrently on side streams forked from (and joined back into)
    // the caller's stream with events, so one batch's tail overlaps the others' work
    static constexpr int NSIDE = 4;
    hipStream_t          side[NSIDE]    = {};
    hipEvent_t           side_ev[NSIDE] = {};
    hipEvent_t           fork_ev        = nullptr;
    // a mixed batch with an fp64 hint runs its routed FP64 pass on aux[k] concurrently with the
    // mixed pass on the batch's stream (side[k] or the caller's)
    hipStream_t aux[NSIDE]      = {};
    hipEvent_t  aux_fork[NSIDE] = {};
    hipEvent_t  aux_join[NSIDE] = {};
    std::vector<void*> owned;
    // this device's copy of the fixed scratch of nnet_compute_noalloc
    uint32_t* scratch_ptr = nullptr;
};

struct nnet_network
{
    nnet_ctx*        ctx = nullptr;
    int              kind = 0, d = 0, nR = 0;
    std::vector<int> rb, pb, ids, nums;
    NetDev           dev{};
    const NetDev*    d_dev = nullptr; // device copy of `dev` (inside `blob`), for the fused launch
    void*            blob  = nullptr;
};

static int fail(nnet_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                                             \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return fail(ctx, NNET_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));          \
    } while (0)

// Byte blob with aligned sub-arrays, copied to the device in one transfer.
struct Blob
{
    std::vector<char> b;
    template<class T>
    size_t add(const T* p, size_t n)
    {
        size_t off = (b.size() + 63) & ~size_t(63);
        b.resize(off + n * sizeof(T));
        if (n) std::memcpy(b.data() + off, p, n * sizeof(T));
        return off;
    }
    template<class T>
    size_t add(const std::vector<T>& v)
    {
        return add(v.data(), v.size());
    }
};

static int build_network(nnet_ctx* ctx, nnet_network* nw)
{
    const int kind = nw->kind, d = nw->d, nR = nw->nR;
    const int nrate = kind == 14 ? 29 : kind == 86 ? 309 : 311;
    if (nR != nrate) return fail(ctx, NNET_E_NETWORK, "reaction count does not match the rate parameterization");
    for (int r = 0; r < nR; ++r)
    {
        if (nw->pb[r] - nw->rb[r] > MAXQ || nw->pb[r] < nw->rb[r] || nw->rb[r + 1] < nw->pb[r])
            return fail(ctx, NNET_E_NETWORK, "malformed reaction list");
        for (int q = nw->rb[r]; q < nw->rb[r + 1]; ++q)
            if (nw->ids[q] < 0 || nw->ids[q] >= d || nw->nums[q] < 1 || nw->nums[q] > 3)
                return fail(ctx, NNET_E_NETWORK, "reaction entry out of range");
    }

    // ---- rate-parameterization tables ----
    const RateTables rt = build_rate_tables(kind);
    const auto &BE0 = rt.BE0, &Z53 = rt.Z53, &fit = rt.fit, &q = rt.q, &choose = rt.choose;
    const auto &ca = rt.ca, &cb = rt.cb, &photo = rt.photo, &scr = rt.scr, &swp = rt.swp, &scl = rt.scl;
    const auto &pe = rt.pe, &pl = rt.pl;
    const int   nfit = rt.nfit, nE = rt.nE, e0 = rt.e0;
    const SpecialIdx si = rt.sp;

    // ---- assembly image (see NetDev) ----
    AsmImage im;
    {
        const std::string err = build_asm_image(kind, d, nR, nw->rb.data(), nw->pb.data(), nw->ids.data(),
                                                nw->nums.data(), im);
        if (!err.empty()) return fail(ctx, NNET_E_NETWORK, err);
    }
    const std::vector<uint32_t>& img     = im.img;
    const std::vector<uint32_t>& lanemap = im.lanemap;
    const int                    nnz     = im.nnz;

    // ---- upload ----
    Blob bl;
    size_t oBE = bl.add(BE0), oZ53 = bl.add(Z53), oFit = bl.add(fit), oQ = bl.add(q), oCh = bl.add(choose),
           oCa = bl.add(ca), oCb = bl.add(cb), oPh = bl.add(photo), oScr = bl.add(scr), oSw = bl.add(swp),
           oScl = bl.add(scl), oPe = bl.add(pe), oPl = bl.add(pl), oElT = bl.add(NNET_ELEC_LOGT, NNET_ELEC_NT),
           oElR = bl.add(NNET_ELEC_LOGRHO, NNET_ELEC_NR),
           oElTab = bl.add(NNET_ELEC_RATE, (size_t)NNET_ELEC_NT * NNET_ELEC_NR * NNET_ELEC_NC), oImg = bl.add(img),
           oLane = bl.add(lanemap), oSelf = bl.add(std::vector<char>(sizeof(NetDev)));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(&nw->blob, bl.b.size()));
    HIP_TRY(ctx, hipMemcpy(nw->blob, bl.b.data(), bl.b.size(), hipMemcpyHostToDevice));
    char*   B = (char*)nw->blob;
    NetDev& n = nw->dev;
    n.net = kind, n.d = d, n.nR = nR, n.nE = nE, n.nfit = nfit, n.fit_e0 = e0, n.rev_div = kind == 14 ? 1 : 0;
    n.nnz = nnz;
    n.sp  = si;
    n.BE0 = (double*)(B + oBE), n.Z53 = (double*)(B + oZ53), n.fit = (double*)(B + oFit), n.q = (double*)(B + oQ);
    n.choose = (double*)(B + oCh), n.ca = (int*)(B + oCa), n.cb = (int*)(B + oCb), n.photo = (int*)(B + oPh);
    n.scr = (int*)(B + oScr), n.swap = (int*)(B + oSw), n.scr_l = (int*)(B + oScl), n.push_e = (int*)(B + oPe);
    n.push_l = (int*)(B + oPl), n.elogT = (double*)(B + oElT), n.elogR = (double*)(B + oElR);
    n.etab = (double*)(B + oElTab), n.BEe = NNET87_ELECTRON_BE_MEV[0] * NNET_MEV_TO_ERG;
    n.img = (uint32_t*)(B + oImg), n.img_words = (int)img.size(), n.oRx = im.oRx, n.oEnt16 = im.oEnt16;
    n.oSpPtr16 = im.oSpPtr16, n.oSpEnt = im.oSpEnt, n.oNzPtr16 = im.oNzPtr16, n.oNzEnt = im.oNzEnt;
    n.oRowPtr16 = im.oRowPtr16, n.oRowEnt = im.oRowEnt, n.oJord16 = im.oJord16, n.oMaPerm16 = im.oMaPerm16, n.lanemap = (const uint4*)(B + oLane);
    HIP_TRY(ctx, hipMemcpy(B + oSelf, &n, sizeof(NetDev), hipMemcpyHostToDevice));
    nw->d_dev = (const NetDev*)(B + oSelf);
    return NNET_OK;
}

// Device buffers of one nnet_compute call: per batch NCTR work-queue counters and BIN_STRIDE
// ordering bins, the dispatch order of the cost-ordered batches and the redo lists of the mixed
// batches (each batch takes its n entries in batch order).
struct CallBufs
{
    unsigned int* queue;
    unsigned int* bins;
    uint32_t*     order;
    uint32_t*     redo;
};
struct Needs
{
    size_t order = 0, redo = 0; // entries
};
static Needs call_needs(const nnet_batch* batches, int nb)
{
    Needs n;
    for (int bi = 0; bi < nb; ++bi)
    {
        const nnet_batch& b = batches[bi];
        if (b.cost_hint || (b.fp64_hint && b.precision == NNET_PRECISION_MIXED)) n.order += b.n;
        if (b.precision == NNET_PRECISION_MIXED) n.redo += b.n;
    }
    return n;
}

// Fixed device scratch of nnet_compute_noalloc, one copy per GPU the library runs on, held until
// the process ends: SCRATCH_CTR_WORDS words for the per-batch work-queue counters and ordering
// bins (NCTR + BIN_STRIDE words per batch), then SCRATCH_LIST_WORDS words for the dispatch-order
// and redo-list entries (one per particle of each cost-ordered batch, one per particle of each
// mixed-precision batch).  64 KiB + 16 MiB.  Host code addresses it only through the pointer that
// hipGetSymbolAddress returns for the current device (scratch_ptr_of), never by its name.
constexpr size_t SCRATCH_CTR_WORDS = 16384, SCRATCH_LIST_WORDS = size_t(4) << 20;
static __device__ __align__(256) uint32_t scratch[SCRATCH_CTR_WORDS + SCRATCH_LIST_WORDS];

// Per-device address of `scratch`, looked up once per device with that device current.
static std::mutex g_scratch_mu;
static std::map<int, uint32_t*> g_scratch_ptr;
static int scratch_ptr_of(int device, uint32_t** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto                        it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end())
    {
        *out = it->second;
        return NNET_OK;
    }
    void* p = nullptr;
    if (hipSetDevice(device) != hipSuccess || hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return NNET_E_HIP;
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    *out                  = static_cast<uint32_t*>(p);
    return NNET_OK;
}

// The autotuner's table (format in nnet_mi355x.h): settings applied over `t`, which is left
// unchanged when the table cannot be read or has a malformed line.
static int tuning_parse(const char* path, Tuning& t, std::string& err)
{
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f)
    {
        err = std::string("cannot open the tuning table ") + (path ? path : "(null)");
        return NNET_E_ARG;
    }
    Tuning r  = t;
    char   line[512];
    int    ln = 0, rc = NNET_OK;
    while (rc == NNET_OK && std::fgets(line, sizeof line, f))
    {
        ++ln;
        if (char* h = std::strchr(line, '#')) *h = 0;
        char key[32], extra[8];
        int  kind, eos, prec, v;
        if (std::sscanf(line, " %31s", key) != 1) continue;
        bool ok = false;
        if (!std::strcmp(key, "ppw") && std::sscanf(line, " ppw %d %d %d %d %7s", &kind, &eos, &prec, &v, extra) == 4 &&
            (kind == 14 || kind == 86 || kind == 87) && (eos == 0 || eos == 1) && (prec == 0 || prec == 1) &&
            (kind == 14 ? (v == 1 || v == 2 || v == 4) : v == 1))
        {
            if (kind == 14) r.ppw14[eos][prec] = v;
            ok = true;
        }
        else if (!std::strcmp(key, "fused") && std::sscanf(line, " fused %d %7s", &v, extra) == 1 && (v == 0 || v == 1))
        {
            r.fused = v;
            ok      = true;
        }
        if (!ok)
        {
            err = std::string("tuning table ") + path + ", line " + std::to_string(ln) +
                  ": expected 'ppw <14|86|87> <eos 0|1> <precision 0|1> <1|2|4 (net14), 1>' or 'fused <0|1>'";
            rc = NNET_E_ARG;
        }
    }
    std::fclose(f);
    if (rc == NNET_OK) t = r;
    return rc;
}

// The table nnet_ctx_create loads: $NNET_MI355X_TUNING (empty: none), else
// data/tuning_gfx950.txt next to the directory of this library when it exists.
static std::string default_tuning_path()
{
    if (const char* e = std::getenv("NNET_MI355X_TUNING")) return e;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&default_tuning_path), &info) && info.dli_fname)
    {
        const std::string lib   = info.dli_fname;
        const size_t      slash = lib.rfind('/');
        const std::string p = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) +
                              "/../data/tuning_gfx950.txt";
        if (FILE* f = std::fopen(p.c_str(), "r"))
        {
            std::fclose(f);
            return p;
        }
    }
    return "";
}

static void tuning_export(const Tuning& t, nnet_tuning* out)
{
    for (int k = 0; k < 3; ++k)
        for (int e = 0; e < 2; ++e)
            for (int p = 0; p < 2; ++p)
                out->ppw[k][e][p] = k == 0 ? t.ppw14[e][p] : 1;
    out->fused = t.fused;
}

// Argument checks of nnet_compute; nothing is allocated or enqueued.  n_live: non-empty batches.
static int check_call(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, int use_drhodt, int& n_live)
{
    if (!c || (nb > 0 && !batches) || nb < 0) return NNET_E_ARG;
    if (!(hydro_dt > 0)) return fail(c, NNET_E_ARG, "hydro_dt must be positive");
    if (use_drhodt != NNET_DRHODT_OFF && use_drhodt != NNET_DRHODT_CPU && use_drhodt != NNET_DRHODT_INTEGRATE)
        return fail(c, NNET_E_ARG, "use_drhodt must be NNET_DRHODT_OFF, _CPU or _INTEGRATE");
    n_live = 0;
    for (int bi = 0; bi < nb; ++bi)
    {
        const nnet_batch& b = batches[bi];
        if (!b.network || b.network->ctx != c) return fail(c, NNET_E_ARG, "batch without a network of this context");
        if (b.n == 0) continue;
        ++n_live;
        if (!b.rho || !b.Y || !b.temp || !b.dt || (use_drhodt != NNET_DRHODT_OFF && !b.rho_m1))
            return fail(c, NNET_E_ARG, "null particle array");
        if (b.precision != NNET_PRECISION_FP64 && b.precision != NNET_PRECISION_MIXED)
            return fail(c, NNET_E_ARG, "unknown precision");
        for (int z = 0; z < b.network->d; ++z)
            if (!b.Y[z]) return fail(c, NNET_E_ARG, "null abundance array");
        if (b.eos == NNET_EOS_HELMHOLTZ)
        {
            if (!c->helm_ok) return fail(c, NNET_E_STATE, "no Helmholtz table loaded");
            if (!b.Z || b.z_dim <= 0 || b.z_dim > b.network->d) return fail(c, NNET_E_ARG, "Helmholtz Z/z_dim invalid");
        }
        else if (b.eos != NNET_EOS_IDEAL)
            return fail(c, NNET_E_ARG, "unknown EOS");
    }
    return NNET_OK;
}

// Is `st` being captured into a HIP graph?  (An unknown status counts as captured.)
static bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus capture = hipStreamCaptureStatusActive;
    if (hipStreamIsCapturing(st, &capture) != hipSuccess)
    {
        (void)hipGetLastError();
        return true;
    }
    return capture != hipStreamCaptureStatusNone;
}

// Checks a call makes before it touches the stream when the stream is being captured.  A
// captured multi-batch call with a mixed-precision batch (a net14 + net86 + net87 hydro loop) ended in a host
// segmentation fault whose cause has not been shown; until it is, that combination is refused
// rather than run (the FP64 multi-batch capture is tested and stays enabled).
static int check_capture(nnet_ctx* c, const nnet_batch* batches, int nb, int n_live)
{
    if (n_live > 1)
        for (int bi = 0; bi < nb; ++bi)
            if (batches[bi].n > 0 && batches[bi].precision == NNET_PRECISION_MIXED)
                return fail(c, NNET_E_STATE, "mixed precision in a multi-batch call under HIP graph capture is disabled");
    return NNET_OK;
}

// Enqueues the kernels of one nnet_compute call on `st` with the given buffers.
// Ordered dispatch of batch bi: longest first by the cost hint; with a routing hint (a mixed
// batch), the particles with a nonzero fp64 hint form the prefix [0, nH) of the order, the rest
// [nH, n).  Returns the order (nullptr: index order, no hint) and points `range` at the device
// bounds {0, nH, n}.
static const uint32_t* enqueue_order(const nnet_batch& b, int bi, const uint32_t* route, const CallBufs& bufs,
                                     size_t& order_off, hipStream_t s, const unsigned int*& range)
{
    range = nullptr;
    if (!b.cost_hint && !route) return nullptr;
    unsigned int* bins  = bufs.bins + (size_t)BIN_STRIDE * bi;
    uint32_t*     order = bufs.order + order_off;
    order_off += b.n;
    const unsigned blocks = (unsigned)std::min<size_t>((b.n + 255) / 256, 2048);
    hipLaunchKernelGGL(nnet_queue_reset, dim3((BIN_STRIDE + 255) / 256), dim3(256), 0, s, bins, BIN_STRIDE);
    hipLaunchKernelGGL(nnet_order_count, dim3(blocks), dim3(256), 0, s, b.cost_hint, route, b.n, bins);
    hipLaunchKernelGGL(nnet_order_offsets, dim3(1), dim3(64), 0, s, bins);
    hipLaunchKernelGGL(nnet_order_scatter, dim3(blocks), dim3(256), 0, s, b.cost_hint, route, b.n, bins, order);
    range = bins + 2 * NKEY; // {0, nH, n}
    return order;
}

// Can the call's batches run as one fused launch?  All live batches share the EOS and the
// precision, there are at most FUSE_PHASES of them, and their abundance pointers and charges fit
// the FUSE_SPECIES-entry arrays of the launch arguments.
static bool fusable(const nnet_batch* batches, int nb)
{
    int live = 0, ys = 0, zs = 0, eos = -1, prec = -1;
    for (int bi = 0; bi < nb; ++bi)
    {
        const nnet_batch& b = batches[bi];
        if (b.n == 0) continue;
        if (eos >= 0 && (b.eos != eos || b.precision != prec)) return false;
        eos = b.eos, prec = b.precision;
        ++live;
        ys += b.network->d;
        zs += b.eos == NNET_EOS_HELMHOLTZ ? b.z_dim : 0;
    }
    return live >= 2 && live <= FUSE_PHASES && ys <= FUSE_SPECIES && zs <= FUSE_SPECIES;
}

// The kernels of a fused call on `st`: per batch the queue reset and the cost ordering, then one
// nnet_solve_fused launch over all batches (mixed: then one FP64 launch over the redo lists).
static int enqueue_fused(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, double previous_dt,
                         int use_drhodt, hipStream_t st, const CallBufs& bufs)
{
    StepParams sp{hydro_dt, previous_dt, use_drhodt, c->max_iter};
    StepParams redo_sp = sp;
    if (sp.use_drhodt == NNET_DRHODT_CPU) redo_sp.use_drhodt = NNET_DRHODT_OFF; // rho_m1 already updated
    // phases in cost order: net87, net86, net14; larger batches first within a network
    std::vector<int> phase;
    for (int bi = 0; bi < nb; ++bi)
        if (batches[bi].n > 0) phase.push_back(bi);
    auto rank = [&](int bi) { return batches[bi].network->kind == 87 ? 0 : batches[bi].network->kind == 86 ? 1 : 2; };
    std::stable_sort(phase.begin(), phase.end(), [&](int x, int y)
                     { return rank(x) != rank(y) ? rank(x) < rank(y) : batches[x].n > batches[y].n; });
    const bool mix = batches[phase[0]].precision == NNET_PRECISION_MIXED;
    const int  eos = batches[phase[0]].eos;
    FusedDev   fd{}, fr{};
    size_t     order_off = 0, redo_off = 0, work = 0;
    int        zoff = 0, yoff = 0;
    fd.nphase = fr.nphase = (int)phase.size();
    for (int k = 0; k < (int)phase.size(); ++k)
    {
        const int         bi  = phase[k];
        const nnet_batch& b   = batches[bi];
        const int         d   = b.network->d;
        unsigned int*     ctr = bufs.queue + (size_t)nnet_ctx::NCTR * bi;
        hipLaunchKernelGGL(nnet_queue_reset, dim3(1), dim3(64), 0, st, ctr, nnet_ctx::NCTR);
        PhaseDev& P = fd.ph[k];
        P.nd        = b.network->d_dev;
        P.kind      = b.network->kind;
        P.eos.kind  = b.eos;
        P.eos.mu    = b.mu;
        P.zoff      = zoff;
        if (b.eos == NNET_EOS_HELMHOLTZ)
        {
            for (int z = 0; z < b.z_dim; ++z)
                fd.Z[zoff + z] = b.Z[z];
            zoff += b.z_dim;
            P.eos.zdim  = b.z_dim;
            P.eos.grid  = c->d_grid;
            P.eos.cells = c->d_cells;
        }
        P.yoff = yoff;
        for (int s = 0; s < d; ++s)
            fd.Y[yoff + s] = b.Y[s];
        yoff += d;
        BatchCore& bc = P.b;
        bc.n = b.n, bc.rho = b.rho, bc.rho_m1 = b.rho_m1, bc.temp = b.temp, bc.dt = b.dt;
        bc.status = b.status, bc.substeps = b.substeps, bc.newton = b.newton, bc.fp64_solves = b.fp64_solves;
        bc.queue = ctr;
        if (mix)
        {
            bc.redo       = bufs.redo + redo_off;
            bc.redo_count = ctr + 4;
            redo_off += b.n;
        }
        const unsigned int* range = nullptr;
        bc.order                  = enqueue_order(b, bi, nullptr, bufs, order_off, st, range);
        // the FP64 pass over the redo list: its own queue, the list as the order, {0, count}
        fr.ph[k]             = P;
        fr.ph[k].b.queue     = ctr + 2;
        fr.ph[k].b.order     = bc.redo;
        fr.ph[k].b.qrange    = ctr + 3;
        fr.ph[k].b.fp64_pass = 1;
        work += P.kind == 14 ? (b.n + 3) / 4 : b.n;
    }
    std::memcpy(fr.Z, fd.Z, sizeof(fd.Z));
    std::memcpy(fr.Y, fd.Y, sizeof(fd.Y));
    auto go = [&](auto kern, const FusedDev& a, const StepParams& s)
    {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, FUSE_NT, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        const size_t wgs = std::min<size_t>(work, (size_t)per_cu * c->n_cu);
        hipLaunchKernelGGL(kern, dim3((unsigned)wgs), dim3(FUSE_NT), 0, st, a, s);
    };
    if (eos == NNET_EOS_IDEAL)
    {
        if (mix)
        {
            go(nnet_solve_fused<NNET_EOS_IDEAL, true>, fd, sp);
            go(nnet_solve_fused<NNET_EOS_IDEAL, false>, fr, redo_sp);
        }
        else
            go(nnet_solve_fused<NNET_EOS_IDEAL, false>, fd, sp);
    }
    else
    {
        if (mix)
        {
            go(nnet_solve_fused<NNET_EOS_HELMHOLTZ, true>, fd, sp);
            go(nnet_solve_fused<NNET_EOS_HELMHOLTZ, false>, fr, redo_sp);
        }
        else
            go(nnet_solve_fused<NNET_EOS_HELMHOLTZ, false>, fd, sp);
    }
    HIP_TRY(c, hipGetLastError());
    return NNET_OK;
}

static int enqueue_compute(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, double previous_dt,
                           int use_drhodt, hipStream_t st, int n_live, const CallBufs& bufs)
{
    // batches of several networks: one fused launch when the context's tuning selects it
    if (n_live > 1 && c->tuning.fused && fusable(batches, nb))
        return enqueue_fused(c, batches, nb, hydro_dt, previous_dt, use_drhodt, st, bufs);
    StepParams sp{hydro_dt, previous_dt, use_drhodt, c->max_iter};
    const bool fork = n_live > 1;
    // the fp64 routing hint forks an auxiliary stream per batch; it is honoured in single-batch
    // calls outside HIP graph capture and ignored otherwise (results stay valid without it)
    const bool allow_route = !fork && !capturing(st);
    if (fork)
    {
        HIP_TRY(c, hipEventRecord(c->fork_ev, st));
        for (int k = 0; k < std::min(n_live, (int)nnet_ctx::NSIDE); ++k)
            HIP_TRY(c, hipStreamWaitEvent(c->side[k], c->fork_ev, 0));
    }
    size_t order_off = 0, redo_off = 0;
    int    live      = 0;
    for (int bi = 0; bi < nb; ++bi)
    {
        const nnet_batch& b = batches[bi];
        if (b.n == 0) continue;
        const hipStream_t bst = fork ? c->side[live++ % nnet_ctx::NSIDE] : st;
        const int         d   = b.network->d;
        EosDev            ed{};
        ed.kind = b.eos;
        ed.mu   = b.mu;
        if (b.eos == NNET_EOS_HELMHOLTZ)
        {
            for (int z = 0; z < b.z_dim; ++z)
                ed.Z[z] = b.Z[z];
            ed.zdim  = b.z_dim;
            ed.grid  = c->d_grid;
            ed.cells = c->d_cells;
        }
        BatchDev bd{};
        bd.n = b.n, bd.rho = b.rho, bd.rho_m1 = b.rho_m1, bd.temp = b.temp, bd.dt = b.dt;
        bd.status = b.status, bd.substeps = b.substeps, bd.newton = b.newton, bd.fp64_solves = b.fp64_solves;
        for (int z = 0; z < d; ++z)
            bd.Y[z] = b.Y[z];
        const NetDev& nd  = b.network->dev;
        unsigned int* ctr = bufs.queue + (size_t)nnet_ctx::NCTR * bi;
        bd.queue          = ctr;
        hipLaunchKernelGGL(nnet_queue_reset, dim3(1), dim3(64), 0, bst, ctr, nnet_ctx::NCTR);
        const bool mix = b.precision == NNET_PRECISION_MIXED;
        if (mix)
        {
            bd.redo       = bufs.redo + redo_off;
            bd.redo_count = ctr + 4;
            redo_off += b.n;
        }
        const uint32_t*     route = (mix && allow_route) ? b.fp64_hint : nullptr;
        const unsigned int* range = nullptr;
        bd.order                  = enqueue_order(b, bi, route, bufs, order_off, bst, range);
        const int         kind = b.network->kind;
        const int         ak   = fork ? (live - 1) % nnet_ctx::NSIDE : 0;
        const hipStream_t ast  = c->aux[ak];
        hipError_t        herr = hipSuccess;
        auto              chk  = [&](hipError_t e)
        {
            if (e != hipSuccess && herr == hipSuccess) herr = e;
        };
        // the solve kernel of the batch's network with the context's particle groups per workgroup
        // (net14: 1, 2 or 4): as many workgroups as are resident at once, at most one group per
        // particle
        auto launch = [&](auto netc, auto eosc, auto mixc, const BatchDev& bdv, const StepParams& spv, hipStream_t s)
        {
            constexpr int  NV = decltype(netc)::value, EV = decltype(eosc)::value;
            constexpr bool MV = decltype(mixc)::value;
            auto           go = [&](auto kern, int ppw)
            {
                const int nt     = Cfg<NV>::NT * ppw;
                int       per_cu = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, 0) != hipSuccess || per_cu < 1)
                    per_cu = 1;
                const size_t wgs = std::min<size_t>((b.n + ppw - 1) / ppw, (size_t)per_cu * c->n_cu);
                hipLaunchKernelGGL(kern, dim3((unsigned)wgs), dim3(nt), 0, s, nd, ed, bdv, spv);
            };
            const int ppw = c->tuning.ppw(NV, EV, MV);
            if constexpr (NV == 14)
            {
                if (ppw == 4) go(nnet_solve_kernel<NV, EV, MV, 4>, 4);
                else if (ppw == 2)
                    go(nnet_solve_kernel<NV, EV, MV, 2>, 2);
                else
                    go(nnet_solve_kernel<NV, EV, MV, 1>, 1);
            }
            else
                go(nnet_solve_kernel<NV, EV, MV, 1>, 1);
        };
        using Fp64  = std::false_type;
        using Mixed = std::true_type;
        // mixed batch: the routed particles [0, nH) go to an FP64 pass on the aux stream while
        // the mixed pass takes [nH, n); the particles the mixed pass gives up go to a last FP64
        // pass over its redo list (rho_m1 already updated by the mixed pass)
        BatchDev routed_bd  = bd;
        routed_bd.queue     = ctr + 1;
        routed_bd.qrange    = range;
        routed_bd.fp64_pass = 1;
        BatchDev mixed_bd   = bd;
        mixed_bd.qrange     = route ? range + 1 : nullptr;
        BatchDev redo_bd    = bd;
        redo_bd.queue       = ctr + 2;
        redo_bd.order       = bd.redo;
        redo_bd.qrange      = ctr + 3; // {0, redo count}
        redo_bd.fp64_pass   = 1;
        StepParams redo_sp  = sp;
        if (sp.use_drhodt == NNET_DRHODT_CPU) redo_sp.use_drhodt = NNET_DRHODT_OFF; // rho_m1 already updated
        auto dispatch       = [&](auto netc, auto eosc)
        {
            if (!mix)
            {
                launch(netc, eosc, Fp64{}, bd, sp, bst);
                return;
            }
            if (route)
            {
                chk(hipEventRecord(c->aux_fork[ak], bst));
                chk(hipStreamWaitEvent(ast, c->aux_fork[ak], 0));
                launch(netc, eosc, Fp64{}, routed_bd, sp, ast);
            }
            launch(netc, eosc, Mixed{}, mixed_bd, sp, bst);
            launch(netc, eosc, Fp64{}, redo_bd, redo_sp, bst);
            if (route)
            {
                chk(hipEventRecord(c->aux_join[ak], ast));
                chk(hipStreamWaitEvent(bst, c->aux_join[ak], 0));
            }
        };
        auto dispatch_eos = [&](auto netc)
        {
            if (b.eos == NNET_EOS_IDEAL) dispatch(netc, std::integral_constant<int, NNET_EOS_IDEAL>{});
            else
                dispatch(netc, std::integral_constant<int, NNET_EOS_HELMHOLTZ>{});
        };
        if (kind == 14) dispatch_eos(std::integral_constant<int, 14>{});
        else if (kind == 86)
            dispatch_eos(std::integral_constant<int, 86>{});
        else
            dispatch_eos(std::integral_constant<int, 87>{});
        HIP_TRY(c, herr);
        HIP_TRY(c, hipGetLastError());
    }
    if (fork)
        for (int k = 0; k < std::min(n_live, (int)nnet_ctx::NSIDE); ++k)
        {
            HIP_TRY(c, hipEventRecord(c->side_ev[k], c->side[k]));
            HIP_TRY(c, hipStreamWaitEvent(st, c->side_ev[k], 0));
        }
    return NNET_OK;
}

extern "C"
{
    int nnet_ctx_create(nnet_ctx** out, int device, const char* helm_table_path)
    {
        if (!out) return NNET_E_ARG;
        *out       = nullptr;
        nnet_ctx* c = new nnet_ctx;
        c->device  = device;
        int ndev   = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        {
            delete c;
            return NNET_E_HIP;
        }
        if (hipSetDevice(device) != hipSuccess ||
            hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        {
            delete c;
            return NNET_E_HIP;
        }
        for (int k = 0; k < nnet_ctx::NSIDE; ++k)
            if (hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&c->side_ev[k], hipEventDisableTiming) != hipSuccess ||
                hipStreamCreateWithFlags(&c->aux[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&c->aux_fork[k], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->aux_join[k], hipEventDisableTiming) != hipSuccess)
            {
                nnet_ctx_destroy(c);
                return NNET_E_HIP;
            }
        if (hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
            scratch_ptr_of(device, &c->scratch_ptr) != NNET_OK)
        {
            nnet_ctx_destroy(c);
            return NNET_E_HIP;
        }
        *out = c;
        const std::string tp = default_tuning_path();
        if (!tp.empty())
        {
            const int rc = tuning_parse(tp.c_str(), c->tuning, c->err);
            if (rc != NNET_OK) return rc;
        }
        if (helm_table_path) return nnet_ctx_load_helmholtz(c, helm_table_path);
        return NNET_OK;
    }

    int nnet_tuning_read(const char* path, nnet_tuning* out)
    {
        if (!path || !out) return NNET_E_ARG;
        Tuning      t;
        std::string err;
        const int   rc = tuning_parse(path, t, err);
        if (rc != NNET_OK) return rc;
        tuning_export(t, out);
        return NNET_OK;
    }

    int nnet_ctx_load_tuning(nnet_ctx* c, const char* path)
    {
        if (!c || !path) return NNET_E_ARG;
        return tuning_parse(path, c->tuning, c->err);
    }

    int nnet_ctx_get_tuning(nnet_ctx* c, nnet_tuning* out)
    {
        if (!c || !out) return NNET_E_ARG;
        tuning_export(c->tuning, out);
        return NNET_OK;
    }

    void nnet_ctx_destroy(nnet_ctx* c)
    {
        if (!c) return;
        (void)hipSetDevice(c->device);
        if (c->d_grid) (void)hipFree(c->d_grid);
        if (c->d_cells) (void)hipFree(c->d_cells);
        if (c->d_cells3) (void)hipFree(c->d_cells3);
        if (c->d_queue) (void)hipFree(c->d_queue);
        if (c->d_bins) (void)hipFree(c->d_bins);
        if (c->d_order) (void)hipFree(c->d_order);
        if (c->d_redo) (void)hipFree(c->d_redo);
        for (int k = 0; k < nnet_ctx::NSIDE; ++k)
        {
            if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
            if (c->side_ev[k]) (void)hipEventDestroy(c->side_ev[k]);
            if (c->aux[k]) (void)hipStreamDestroy(c->aux[k]);
            if (c->aux_fork[k]) (void)hipEventDestroy(c->aux_fork[k]);
            if (c->aux_join[k]) (void)hipEventDestroy(c->aux_join[k]);
        }
        if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
        for (void* p : c->owned)
            (void)hipFree(p);
        delete c;
    }

    const char* nnet_last_error(const nnet_ctx* c) { return c ? c->err.c_str() : "null context"; }

    int nnet_ctx_set_max_iterations(nnet_ctx* c, long m)
    {
        if (!c || m < 0) return NNET_E_ARG;
        c->max_iter = m;
        return NNET_OK;
    }

    int nnet_solve_slots(nnet_ctx* c, int rates_kind, int eos, int precision, int* slots)
    {
        if (!c || !slots) return NNET_E_ARG;
        if (!(rates_kind == 14 || rates_kind == 86 || rates_kind == 87) ||
            !(eos == NNET_EOS_IDEAL || eos == NNET_EOS_HELMHOLTZ) ||
            !(precision == NNET_PRECISION_FP64 || precision == NNET_PRECISION_MIXED))
            return fail(c, NNET_E_ARG, "nnet_solve_slots: unknown network, EOS or precision");
        HIP_TRY(c, hipSetDevice(c->device));
        int        per_cu = 0;
        hipError_t e      = hipSuccess;
        const bool mixed  = precision == NNET_PRECISION_MIXED;
        const int  ppw    = c->tuning.ppw(rates_kind, eos, mixed);
        auto       occ    = [&](auto kern, int nt) { e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, 0); };
        auto       with   = [&](auto netc, auto eosc, auto mixc)
        {
            constexpr int  NV = decltype(netc)::value, EV = decltype(eosc)::value;
            constexpr bool MV = decltype(mixc)::value;
            if constexpr (NV == 14)
            {
                if (ppw == 4) occ(nnet_solve_kernel<NV, EV, MV, 4>, Cfg<NV>::NT * 4);
                else if (ppw == 2)
                    occ(nnet_solve_kernel<NV, EV, MV, 2>, Cfg<NV>::NT * 2);
                else
                    occ(nnet_solve_kernel<NV, EV, MV, 1>, Cfg<NV>::NT);
            }
            else
                occ(nnet_solve_kernel<NV, EV, MV, 1>, Cfg<NV>::NT);
        };
        auto pick = [&](auto netc, auto eosc)
        {
            if (mixed) with(netc, eosc, std::true_type{});
            else
                with(netc, eosc, std::false_type{});
        };
        auto by_eos = [&](auto netc)
        {
            if (eos == NNET_EOS_IDEAL) pick(netc, std::integral_constant<int, NNET_EOS_IDEAL>{});
            else
                pick(netc, std::integral_constant<int, NNET_EOS_HELMHOLTZ>{});
        };
        if (rates_kind == 14) by_eos(std::integral_constant<int, 14>{});
        else if (rates_kind == 86)
            by_eos(std::integral_constant<int, 86>{});
        else
            by_eos(std::integral_constant<int, 87>{});
        HIP_TRY(c, e);
        *slots = std::max(per_cu, 1) * ppw * c->n_cu;
        return NNET_OK;
    }

    int nnet_ctx_set_particles_per_workgroup(nnet_ctx* c, int rates_kind, int eos, int precision, int ppw)
    {
        if (!c) return NNET_E_ARG;
        if (!(rates_kind == 14 || rates_kind == 86 || rates_kind == 87) ||
            !(eos == NNET_EOS_IDEAL || eos == NNET_EOS_HELMHOLTZ) ||
            !(precision == NNET_PRECISION_FP64 || precision == NNET_PRECISION_MIXED))
            return fail(c, NNET_E_ARG, "particles per workgroup: unknown network, EOS or precision");
        if (rates_kind == 14 ? !(ppw == 1 || ppw == 2 || ppw == 4) : ppw != 1)
            return fail(c, NNET_E_ARG, "particles per workgroup: 1, 2 or 4 for net14, 1 for net86 and net87");
        if (rates_kind == 14) c->tuning.ppw14[eos == NNET_EOS_HELMHOLTZ][precision == NNET_PRECISION_MIXED] = ppw;
        return NNET_OK;
    }

    int nnet_ctx_set_fused_launch(nnet_ctx* c, int fused)
    {
        if (!c || (fused != 0 && fused != 1)) return NNET_E_ARG;
        c->tuning.fused = fused;
        return NNET_OK;
    }

    int nnet_ctx_get_fused_launch(nnet_ctx* c, int* fused)
    {
        if (!c || !fused) return NNET_E_ARG;
        *fused = c->tuning.fused;
        return NNET_OK;
    }

    int nnet_ctx_get_particles_per_workgroup(nnet_ctx* c, int rates_kind, int eos, int precision, int* ppw)
    {
        if (!c || !ppw) return NNET_E_ARG;
        if (!(rates_kind == 14 || rates_kind == 86 || rates_kind == 87) ||
            !(eos == NNET_EOS_IDEAL || eos == NNET_EOS_HELMHOLTZ) ||
            !(precision == NNET_PRECISION_FP64 || precision == NNET_PRECISION_MIXED))
            return fail(c, NNET_E_ARG, "particles per workgroup: unknown network, EOS or precision");
        *ppw = c->tuning.ppw(rates_kind, eos, precision == NNET_PRECISION_MIXED);
        return NNET_OK;
    }

    int nnet_helm_table_read(const char* path, double* out)
    {
        if (!path || !out) return NNET_E_ARG;
        const size_t cells = (size_t)HELM_IMAX * HELM_JMAX, total = 21 * cells;
        FILE*        f     = std::fopen(path, "rb");
        if (!f) return NNET_E_TABLE;
        std::string buf;
        {
            char   chunk[1 << 16];
            size_t got;
            while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0)
                buf.append(chunk, got);
        }
        std::fclose(f);
        // the binary layout: 21 arrays [density][temperature], nothing else
        if (buf.size() == total * sizeof(double))
        {
            std::memcpy(out, buf.data(), buf.size());
            return NNET_OK;
        }
        // the reference's text table (helmholtz.cpp:89-137): per temperature j and density i the
        // nine free-energy values, then the dpdf, ef and xf blocks of four, whitespace-separated;
        // optionally wrapped in STR( ... ) as the reference #includes it (helmholtz.cpp:34-35, 92-94)
        const char* q   = buf.c_str();
        const char* end = q + buf.size();
        while (q < end && std::isspace((unsigned char)*q))
            ++q;
        if (end - q >= 4 && std::strncmp(q, "STR(", 4) == 0)
        {
            q += 4;
            while (end > q && std::isspace((unsigned char)end[-1]))
                --end;
            if (end == q || end[-1] != ')') return NNET_E_TABLE;
            --end;
        }
        const size_t q_off = (size_t)(q - buf.c_str());
        buf.resize((size_t)(end - buf.c_str())); // the number stream ends at the closing parenthesis
        q = buf.c_str() + q_off;
        auto next = [&](double& v) -> bool
        {
            char* e = nullptr;
            v       = std::strtod(q, &e);
            if (e == q) return false;
            q = e;
            return true;
        };
        auto at = [&](int k, int i, int j) -> double& { return out[(size_t)k * cells + (size_t)i * HELM_JMAX + j]; };
        for (int j = 0; j < HELM_JMAX; ++j)
            for (int i = 0; i < HELM_IMAX; ++i)
                for (int k = 0; k < 9; ++k)
                    if (!next(at(k, i, j))) return NNET_E_TABLE;
        for (int blk = 0; blk < 3; ++blk)
            for (int j = 0; j < HELM_JMAX; ++j)
                for (int i = 0; i < HELM_IMAX; ++i)
                    for (int k = 0; k < 4; ++k)
                        if (!next(at(9 + 4 * blk + k, i, j))) return NNET_E_TABLE;
        while (*q && std::isspace((unsigned char)*q))
            ++q;
        return *q ? NNET_E_TABLE : NNET_OK; // nothing may follow the last value
    }

    int nnet_ctx_load_helmholtz(nnet_ctx* c, const char* path)
    {
        if (!c || !path) return NNET_E_ARG;
        const size_t        cells = (size_t)HELM_IMAX * HELM_JMAX;
        std::vector<double> arr((size_t)21 * cells);
        const int           rc = nnet_helm_table_read(path, arr.data());
        if (rc != NNET_OK)
            return fail(c, rc, std::string("cannot read Helmholtz table ") + path +
                                   " (binary 21 x 541 x 201 doubles, or the reference's text table)");
        // grid (helmholtz.cpp:101-169)
        HelmGrid g{};
        for (int i = 0; i < HELM_IMAX; ++i)
            g.d[i] = std::pow((double)10., helmc::dlo + i * helmc::dstp);
        for (int j = 0; j < HELM_JMAX; ++j)
            g.t_[j] = std::pow((double)10., helmc::tlo + j * helmc::tstp);
        for (int j = 0; j < HELM_JMAX - 1; ++j)
        {
            const double dth = g.t_[j + 1] - g.t_[j], dt2 = dth * dth;
            g.dt_sav[j] = dth, g.dt2_sav[j] = dt2, g.dti_sav[j] = 1. / dth, g.dt2i_sav[j] = 1. / dt2;
        }
        for (int i = 0; i < HELM_IMAX - 1; ++i)
        {
            const double dd = g.d[i + 1] - g.d[i], dd2 = dd * dd;
            g.dd_sav[i] = dd, g.dd2_sav[i] = dd2, g.ddi_sav[i] = 1. / dd;
        }
        // 4-corner stencil per cell, coefficient order of helmholtz.cpp:431-466
        const int           ks[9] = {0, 2, 4, 1, 3, 5, 6, 7, 8}; // f, ft, ftt, fd, fdd, fdt, fddt, fdtt, fddtt
        std::vector<double> st(cells * HELM_CELL, 0.);
        for (int i = 0; i + 1 < HELM_IMAX; ++i)
            for (int j = 0; j + 1 < HELM_JMAX; ++j)
            {
                double* o = &st[((size_t)i * HELM_JMAX + j) * HELM_CELL];
                for (int qq = 0; qq < 9; ++qq)
                {
                    const double* a = &arr[(size_t)ks[qq] * cells];
                    o[4 * qq + 0]   = a[(size_t)(i + 0) * HELM_JMAX + (j + 0)];
                    o[4 * qq + 1]   = a[(size_t)(i + 1) * HELM_JMAX + (j + 0)];
                    o[4 * qq + 2]   = a[(size_t)(i + 0) * HELM_JMAX + (j + 1)];
                    o[4 * qq + 3]   = a[(size_t)(i + 1) * HELM_JMAX + (j + 1)];
                }
            }
        // bicubic dP/drho stencil per cell, coefficient order of helmholtz.cpp:587-602
        const int           ks3[4] = {9, 11, 10, 12}; // dpdf, dpdft, dpdfd, dpdfdt
        std::vector<double> st3(cells * HELM_CELL3, 0.);
        for (int i = 0; i + 1 < HELM_IMAX; ++i)
            for (int j = 0; j + 1 < HELM_JMAX; ++j)
            {
                double* o = &st3[((size_t)i * HELM_JMAX + j) * HELM_CELL3];
                for (int qq = 0; qq < 4; ++qq)
                {
                    const double* a = &arr[(size_t)ks3[qq] * cells];
                    o[4 * qq + 0]   = a[(size_t)(i + 0) * HELM_JMAX + (j + 0)];
                    o[4 * qq + 1]   = a[(size_t)(i + 1) * HELM_JMAX + (j + 0)];
                    o[4 * qq + 2]   = a[(size_t)(i + 0) * HELM_JMAX + (j + 1)];
                    o[4 * qq + 3]   = a[(size_t)(i + 1) * HELM_JMAX + (j + 1)];
                }
            }
        HIP_TRY(c, hipSetDevice(c->device));
        if (!c->d_grid) HIP_TRY(c, hipMalloc(&c->d_grid, sizeof(HelmGrid)));
        if (!c->d_cells) HIP_TRY(c, hipMalloc(&c->d_cells, st.size() * sizeof(double)));
        if (!c->d_cells3) HIP_TRY(c, hipMalloc(&c->d_cells3, st3.size() * sizeof(double)));
        HIP_TRY(c, hipMemcpy(c->d_grid, &g, sizeof(HelmGrid), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->d_cells, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->d_cells3, st3.data(), st3.size() * sizeof(double), hipMemcpyHostToDevice));
        c->helm_ok = true;
        return NNET_OK;
    }

    int nnet_network_create(nnet_ctx* c, int kind, int n_species, int n_reactions, const int* reactant_begin,
                            const int* product_begin, const int* species_id, const int* count, nnet_network** out)
    {
        if (!c || !out || !reactant_begin || !product_begin || !species_id || !count) return NNET_E_ARG;
        *out = nullptr;
        if (!(kind == 14 || kind == 86 || kind == 87)) return fail(c, NNET_E_ARG, "unknown rate parameterization");
        const int dexp = kind == 14 ? 14 : kind == 86 ? 86 : 87;
        if (n_species != dexp) return fail(c, NNET_E_NETWORK, "species count does not match the rate parameterization");
        nnet_network* nw = new nnet_network;
        nw->ctx = c, nw->kind = kind, nw->d = n_species, nw->nR = n_reactions;
        nw->rb.assign(reactant_begin, reactant_begin + n_reactions + 1);
        nw->pb.assign(product_begin, product_begin + n_reactions);
        nw->ids.assign(species_id, species_id + nw->rb.back());
        nw->nums.assign(count, count + nw->rb.back());
        const int rc = build_network(c, nw);
        if (rc != NNET_OK)
        {
            if (nw->blob) (void)hipFree(nw->blob);
            delete nw;
            return rc;
        }
        *out = nw;
        return NNET_OK;
    }

    int nnet_network_builtin(nnet_ctx* c, int kind, nnet_network** out)
    {
        if (!c || !out) return NNET_E_ARG;
        if (!(kind == 14 || kind == 86 || kind == 87)) return fail(c, NNET_E_ARG, "unknown rate parameterization");
        HostList L = builtin_list(kind);
        if (L.pb.empty()) return fail(c, NNET_E_NETWORK, "could not build the reaction list");
        const int d = kind == 14 ? 14 : kind == 86 ? 86 : 87;
        return nnet_network_create(c, kind, d, (int)L.pb.size(), L.rb.data(), L.pb.data(), L.ids.data(),
                                   L.nums.data(), out);
    }

    void nnet_network_destroy(nnet_network* nw)
    {
        if (!nw) return;
        if (nw->blob) (void)hipFree(nw->blob);
        delete nw;
    }

    int nnet_network_reactions(const nnet_network* nw, int* n_reactions, int* n_entries, int* reactant_begin,
                               int* product_begin, int* species_id, int* count)
    {
        if (!nw) return NNET_E_ARG;
        if (n_reactions) *n_reactions = nw->nR;
        if (n_entries) *n_entries = (int)nw->ids.size();
        if (reactant_begin) std::memcpy(reactant_begin, nw->rb.data(), nw->rb.size() * sizeof(int));
        if (product_begin) std::memcpy(product_begin, nw->pb.data(), nw->pb.size() * sizeof(int));
        if (species_id) std::memcpy(species_id, nw->ids.data(), nw->ids.size() * sizeof(int));
        if (count) std::memcpy(count, nw->nums.data(), nw->nums.size() * sizeof(int));
        return NNET_OK;
    }

    int nnet_compute(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, double previous_dt,
                     int use_drhodt, void* stream)
    {
        int n_live = 0;
        int rc     = check_call(c, batches, nb, hydro_dt, use_drhodt, n_live);
        if (rc != NNET_OK) return rc;
        HIP_TRY(c, hipSetDevice(c->device));
        hipStream_t st   = (hipStream_t)stream;
        const Needs need = call_needs(batches, nb);
        const bool  grow = nb > c->n_queue || need.order > c->n_order || (need.order > 0 && nb > c->n_bins) ||
                          need.redo > c->n_redo;
        if (capturing(st))
        {
            // nothing that would invalidate the caller's capture (a synchronisation, an allocation)
            if ((rc = check_capture(c, batches, nb, n_live)) != NNET_OK) return rc;
            if (grow)
                return fail(c, NNET_E_STATE, "the library's buffers must grow for this call: make the same call once "
                                             "outside HIP graph capture first");
        }
        if (nb > c->n_queue)
        {
            // grows only in calls with more batches than before (a capture-safe caller keeps nb fixed)
            HIP_TRY(c, hipStreamSynchronize(st));
            if (c->d_queue) HIP_TRY(c, hipFree(c->d_queue));
            c->d_queue = nullptr;
            HIP_TRY(c, hipMalloc(&c->d_queue, sizeof(unsigned int) * nnet_ctx::NCTR * std::max(nb, 8)));
            c->n_queue = std::max(nb, 8);
        }
        // cost-ordered dispatch buffers: grow only when a call needs more than before
        if (need.order > c->n_order || (need.order > 0 && nb > c->n_bins))
        {
            HIP_TRY(c, hipStreamSynchronize(st));
            if (need.order > c->n_order)
            {
                if (c->d_order) HIP_TRY(c, hipFree(c->d_order));
                c->d_order = nullptr;
                HIP_TRY(c, hipMalloc(&c->d_order, sizeof(uint32_t) * need.order));
                c->n_order = need.order;
            }
            if (nb > c->n_bins)
            {
                if (c->d_bins) HIP_TRY(c, hipFree(c->d_bins));
                c->d_bins = nullptr;
                HIP_TRY(c, hipMalloc(&c->d_bins, sizeof(unsigned int) * BIN_STRIDE * std::max(nb, 8)));
                c->n_bins = std::max(nb, 8);
            }
        }
        // redo lists of the mixed batches: grow only when a call needs more than before
        if (need.redo > c->n_redo)
        {
            HIP_TRY(c, hipStreamSynchronize(st));
            if (c->d_redo) HIP_TRY(c, hipFree(c->d_redo));
            c->d_redo = nullptr;
            HIP_TRY(c, hipMalloc(&c->d_redo, sizeof(uint32_t) * need.redo));
            c->n_redo = need.redo;
        }
        return enqueue_compute(c, batches, nb, hydro_dt, previous_dt, use_drhodt, st, n_live,
                               CallBufs{c->d_queue, c->d_bins, c->d_order, c->d_redo});
    }

    int nnet_compute_sync(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, double previous_dt,
                          int use_drhodt)
    {
        int rc = nnet_compute(c, batches, nb, hydro_dt, previous_dt, use_drhodt, nullptr);
        if (rc != NNET_OK) return rc;
        HIP_TRY(c, hipDeviceSynchronize());
        return NNET_OK;
    }

    int nnet_compute_noalloc(nnet_ctx* c, const nnet_batch* batches, int nb, double hydro_dt, double previous_dt,
                             int use_drhodt, void* stream)
    {
        int n_live = 0;
        int rc     = check_call(c, batches, nb, hydro_dt, use_drhodt, n_live);
        if (rc != NNET_OK) return rc;
        if (!c->scratch_ptr) return fail(c, NNET_E_STATE, "nnet_compute_noalloc: no fixed scratch on this device");
        const Needs  need = call_needs(batches, nb);
        const size_t ctr  = (size_t)(nnet_ctx::NCTR + BIN_STRIDE) * (size_t)nb;
        if (ctr > SCRATCH_CTR_WORDS)
            return fail(c, NNET_E_ARG, "nnet_compute_noalloc: at most " +
                                           std::to_string(SCRATCH_CTR_WORDS / (nnet_ctx::NCTR + BIN_STRIDE)) +
                                           " batches per call");
        if (need.order + need.redo > SCRATCH_LIST_WORDS)
            return fail(c, NNET_E_ARG, "nnet_compute_noalloc: at most " + std::to_string(SCRATCH_LIST_WORDS) +
                                           " particles per call counting each cost-hinted and each mixed-precision "
                                           "batch once (both: twice)");
        HIP_TRY(c, hipSetDevice(c->device));
        if (capturing((hipStream_t)stream) && (rc = check_capture(c, batches, nb, n_live)) != NNET_OK) return rc;
        uint32_t* base = c->scratch_ptr;
        CallBufs  bufs{base, base + (size_t)nnet_ctx::NCTR * nb, base + SCRATCH_CTR_WORDS,
                      base + SCRATCH_CTR_WORDS + need.order};
        return enqueue_compute(c, batches, nb, hydro_dt, previous_dt, use_drhodt, (hipStream_t)stream, n_live, bufs);
    }

    int nnet_helm_eos(nnet_ctx* c, size_t n, int dimension, const double* Z, const double* temp, const double* rho,
                      const double* const* Y, double* u, double* cv, double* p, double* cs, double* dpdT, void* stream)
    {
        if (!c) return NNET_E_ARG;
        if (dimension < 1 || dimension > HELM_EOS_MAX_DIM)
            return fail(c, NNET_E_ARG, "nnet_helm_eos: dimension must be in [1, " + std::to_string(HELM_EOS_MAX_DIM) + "]");
        if (!Z || !Y) return fail(c, NNET_E_ARG, "nnet_helm_eos: Z and Y are required");
        if (!c->helm_ok) return fail(c, NNET_E_STATE, "no Helmholtz table loaded");
        if (n == 0) return NNET_OK;
        if (!temp || !rho) return fail(c, NNET_E_ARG, "nnet_helm_eos: temp and rho are required");
        if (n > (size_t)0x7fffffff * 256) return fail(c, NNET_E_ARG, "nnet_helm_eos: too many particles for one call");
        HelmEosArgs a{};
        a.n = n, a.dim = dimension, a.temp = temp, a.rho = rho;
        a.u = u, a.cv = cv, a.p = p, a.c = cs, a.dpdT = dpdT;
        for (int j = 0; j < dimension; ++j)
        {
            if (!Y[j]) return fail(c, NNET_E_ARG, "nnet_helm_eos: null abundance array");
            a.Y[j] = Y[j];
            a.Z[j] = Z[j];
        }
        HIP_TRY(c, hipSetDevice(c->device));
        hipLaunchKernelGGL(nnet_helm_eos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, a, (const HelmGrid*)c->d_grid, (const double*)c->d_cells,
                           (const double*)c->d_cells3);
        HIP_TRY(c, hipGetLastError());
        return NNET_OK;
    }

    int nnet_helm_eos_sync(nnet_ctx* c, size_t n, int dimension, const double* Z, const double* temp,
                           const double* rho, const double* const* Y, double* u, double* cv, double* p, double* cs,
                           double* dpdT)
    {
        int rc = nnet_helm_eos(c, n, dimension, Z, temp, rho, Y, u, cv, p, cs, dpdT, nullptr);
        if (rc != NNET_OK) return rc;
        HIP_TRY(c, hipDeviceSynchronize());
        return NNET_OK;
    }
}
#endif // NNET_SOLVE_NET
