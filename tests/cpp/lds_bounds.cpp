// Host-side check of every LDS index the solve kernels form (nnet_solver.hip), for the net14,
// net86 and net87 configurations in FP64 and mixed precision.  The kernels address LDS through
// the constexpr layout (Lds<>), the LU index functions (LuIdx<>) and the network images built by
// build_asm_image / build_rate_tables; this program includes those same definitions, walks the
// kernels' loop nests over every lane, block and register slot, and checks each index against
// the array it addresses and the LDS allocation.  Where a load sits under a lane condition the
// kernels select the index before the load (`lds[cond ? i : 0]`), so the address is valid for
// every lane whether or not the compiler issues the load before the select; the remaining
// operand loads of that form (the MFMA A/B operands) are checked with the condition ignored.
//
// Exit status 0 when every index is in range; otherwise each violation is printed.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../sph-exa-nnet_amd/csrc/nnet_image.hpp"

using namespace nnet_amd;

namespace
{
int  n_checks = 0, n_fail = 0;
void check(bool ok, const std::string& what, long idx, long lo, long hi)
{
    ++n_checks;
    if (!ok && n_fail++ < 200) std::printf("FAIL %s: index %ld outside [%ld, %ld)\n", what.c_str(), idx, lo, hi);
}
#define IN(name, idx, lo, hi) check((long)(idx) >= (long)(lo) && (long)(idx) < (long)(hi), name, idx, lo, hi)

template<int NET, bool MIX>
void run()
{
    using C = Cfg<NET>;
    using L = Lds<NET, MIX>;
    using X = LuIdx<NET>;
    constexpr int D = C::D, N = C::N, NT = C::NT, WR = C::WR, TI = C::TT / C::WR, NPAD = X::NPAD, NA = (N + 63) / 64;
    const std::string tag = "net" + std::to_string(NET) + (MIX ? "/mixed " : "/fp64 ");
    std::printf("%s LDS %d doubles + %d image words = %d bytes\n", tag.c_str(), L::total, C::IMG,
                L::total * 8 + C::IMG * 4 + 8);

    // ---- allocation: the kernel's LDS fits the CU, twice for the 256-thread kernels ----
    const long bytes = (long)L::total * 8 + (long)C::IMG * 4 + 8, per_cu = NT == 256 ? 2 : 1;
    check(bytes * per_cu <= 160 * 1024, tag + "LDS per CU (bytes)", bytes * per_cu, 0, 160 * 1024 + 1);

    // ---- network data ----
    HostList         hl = builtin_list(NET);
    AsmImage         im;
    const int        nR = (int)hl.pb.size();
    const std::string err = build_asm_image(NET, D, nR, hl.rb.data(), hl.pb.data(), hl.ids.data(), hl.nums.data(), im);
    check(err.empty(), tag + "image builds (" + err + ")", 0, 0, 1);
    const RateTables rt = build_rate_tables(NET);
    const int        nnz = im.nnz, W = (int)im.img.size();
    IN(tag + "image words", W, 0, C::IMG + 1);
    IN(tag + "nnz", nnz, 0, C::NNZ);
    IN(tag + "nR", nR, 0, C::NR + 1);
    IN(tag + "nE", rt.nE, 0, C::NE + 1);
    const uint16_t* h16 = reinterpret_cast<const uint16_t*>(im.img.data());
    auto            w32 = [&](int o) { return im.img.at(o); };

    // image sub-arrays (words / halfwords) inside the image; their entries index what they claim
    const int E = hl.rb[nR];
    for (int r = 0; r <= nR; ++r)
        IN(tag + "rx", im.oRx + r, 0, W);
    for (int r = 0; r < nR; ++r)
    {
        const int q0 = w32(im.oRx + r) & 0xffff, q1 = w32(im.oRx + r) >> 16;
        check(q1 - q0 <= MAXQ && q0 <= q1, tag + "reactants per reaction", q1 - q0, 0, MAXQ + 1);
        for (int q = q0; q < q1; ++q)
        {
            IN(tag + "ent16", im.oEnt16 + q, 0, 2 * W);
            IN(tag + "ent16 species", h16[im.oEnt16 + q] & 0xff, 0, D);
        }
    }
    for (int s = 0; s <= D; ++s)
        IN(tag + "sp_ptr16", im.oSpPtr16 + s, 0, 2 * W);
    for (int e = h16[im.oSpPtr16]; e < h16[im.oSpPtr16 + D]; ++e)
    {
        IN(tag + "sp_ent", im.oSpEnt + e, 0, W);
        IN(tag + "sp_ent reaction", w32(im.oSpEnt + e) & 0xffff, 0, nR);
    }
    for (int e = 0; e <= nnz; ++e)
        IN(tag + "nz_ptr16", im.oNzPtr16 + e, 0, 2 * W);
    for (int k = h16[im.oNzPtr16]; k < h16[im.oNzPtr16 + nnz]; ++k)
    {
        IN(tag + "nz_ent", im.oNzEnt + k, 0, W);
        const uint32_t w = w32(im.oNzEnt + k);
        IN(tag + "nz_ent reaction", w & 0xfff, 0, nR);
        IN(tag + "nz_ent reactant position", (w >> 12) & 0xf, 0, MAXQ);
    }
    for (int a = 0; a <= D; ++a)
        IN(tag + "row_ptr16", im.oRowPtr16 + a, 0, 2 * W);
    for (int e = h16[im.oRowPtr16]; e < h16[im.oRowPtr16 + D]; ++e)
    {
        IN(tag + "row_ent", im.oRowEnt + e, 0, W);
        IN(tag + "row_ent non-zero", w32(im.oRowEnt + e) & 0xffff, 0, nnz);
        IN(tag + "row_ent column", w32(im.oRowEnt + e) >> 16, 0, D);
    }
    for (int k = 0; k < nnz; ++k)
    {
        IN(tag + "jord16", im.oJord16 + k, 0, 2 * W);
        IN(tag + "jord16 non-zero", h16[im.oJord16 + k], 0, nnz);
    }
    // Jacobian phase: every non-zero slot is summed by exactly one (lane, pass) of the kernel's loop
    {
        std::vector<int> seen(nnz, 0);
        for (int tid = 0; tid < NT; ++tid)
            for (int pass = 0;; ++pass)
            {
                const int k = jac_slot(tid, pass, NT, nnz);
                if (k >= nnz) break;
                IN(tag + "Jacobian slot", k, 0, nnz);
                if (k >= 0) ++seen[k];
                check(pass < 64, tag + "Jacobian passes end", pass, 0, 64);
                if (pass >= 64) break;
            }
        for (int k = 0; k < nnz; ++k)
            check(seen[k] == 1, tag + "Jacobian slot summed once", seen[k], 1, 2);
    }
    // mass-action order: a permutation of the reactions (each computed by exactly one slot)
    {
        std::vector<int> seen(nR, 0);
        for (int k = 0; k < nR; ++k)
        {
            IN(tag + "maperm16", im.oMaPerm16 + k, 0, 2 * W);
            const int r = h16[im.oMaPerm16 + k];
            IN(tag + "maperm16 reaction", r, 0, nR);
            if (r >= 0 && r < nR) ++seen[r];
        }
        for (int r = 0; r < nR; ++r)
            check(seen[r] == 1, tag + "maperm16 reaction computed once", seen[r], 1, 2);
    }
    (void)E;

    // ---- the LDS arrays of the kernel (offsets in doubles) ----
    const int total = L::total;
    // head
    for (int s = 0; s < (NET == 14 ? 14 : 86); ++s)
        IN(tag + "sMu[s]", L::oMu + s, L::oMu, L::oEff);
    for (int i = 0; i < rt.nE; ++i)
        for (int k = 0; k < 4; ++k)
        {
            const int e = rt.scr[4 * i + k];
            IN(tag + "sMu[scr]", L::oMu + (k < 3 || e >= 0 ? e : 0), L::oMu, L::oEff);
        }
    for (int f = 0; f < rt.nfit; ++f)
        IN(tag + "sEff[fit_e0 + f]", L::oEff + rt.e0 + f, L::oEff, L::oDeff);
    const int sp[] = {rt.sp.r3a_e, rt.sp.r3a_l, rt.sp.r24, rt.sp.r1216, rt.sp.r32, rt.sp.rcag_e, rt.sp.rcag_l,
                      rt.sp.roag_e, rt.sp.roag_l};
    for (int e : sp)
        IN(tag + "hand-fitted slot", L::oEff + e, L::oEff, L::oDeff);
    for (int i = 0; i < rt.nE; ++i)
    {
        if (rt.pe[i] >= 0) IN(tag + "sRate[push_e]", L::oRate + rt.pe[i], L::oRate, L::oDrate);
        if (rt.pl[i] >= 0) IN(tag + "sRate[push_l]", L::oRate + rt.pl[i], L::oRate, L::oDrate);
    }
    if (NET == 87)
        for (int r : {309, 310})
            IN(tag + "sRate[electron capture]", L::oRate + r, L::oRate, L::oDrate);
    for (int r = 0; r < nR; ++r)
    {
        IN(tag + "sVal[r]", L::oVal + r, L::oVal, L::oDval);
        IN(tag + "sJok[r]", L::oJok + r, L::oJok, L::headEnd);
        const int q0 = w32(im.oRx + r) & 0xffff, q1 = w32(im.oRx + r) >> 16;
        for (int q = q0; q < q1; ++q)
            IN(tag + "sTr[r*MAXQ+q]", L::oTr + r * MAXQ + (q - q0), L::oTr, L::oJok);
    }
    // tail
    for (int s = 0; s < D; ++s)
    {
        IN(tag + "sBE[s]", L::oBE + s, L::oBE, L::oDYdot);
        IN(tag + "sDYd[s]", L::oDYdot + s, L::oDYdot, L::oDYdT);
        IN(tag + "sDYT[s]", L::oDYdT + s, L::oDYdT, L::oCorr);
        IN(tag + "sCorr[s]", L::oCorr + s, L::oCorr, L::oEv);
        IN(tag + "sY[s]", L::oY + s, L::oY, L::oYb);
        IN(tag + "sYb[s]", L::oYb + s, L::oYb, L::oX);
    }
    for (int c = 0; c < 12; ++c)
        IN(tag + "sEv[c]", L::oEv + c, L::oEv, L::oEos);
    for (int c = 0; c < 2; ++c)
        IN(tag + "sEos[c]", L::oEos + c, L::oEos, L::oJv);
    for (int e = 0; e < nnz; ++e)
        IN(tag + "sJv[e]", L::oJv + e, L::oJv, L::phaseEnd);
    // the hand-fitted groups' exp / pow values, staged in sJv during the rate phase
    for (int k = 0; k < spx::N1 + spx::NE2; ++k)
        IN(tag + "hand-fitted value sJv[k]", L::oJv + k, L::oJv, L::oJv + C::NNZ);
    for (int r = 0; r <= N; ++r)
        IN(tag + "sX[r]", L::oX + r, L::oX, L::oRow0);
    if (MIX)
        for (int c = 0; c <= N; ++c)
            IN(tag + "sRow0[c]", L::oRow0 + c, L::oRow0, L::oPan);

    // ---- matrix fill: register slots, non-zero ids, entry() operands ----
    // Jacobian block for every slot (index selected before the load), then the column-0,
    // column-N and row-0 patches; every matrix entry (r, c), r < N, c <= N, must be written by
    // exactly one of: the block (1 <= r, c < N) or one patch.
    std::vector<int> written((size_t)N * (N + 1), 0);
    auto             mark = [&](int r, int c) { if (r >= 0 && r < N && c >= 0 && c <= N) ++written[(size_t)r * (N + 1) + c]; };
    constexpr int    JN   = N >> 4;
    for (int tid = 0; tid < NT; ++tid)
    {
        const int      lr = (tid & 63) >> 4, lc = tid & 15, wr = (tid >> 6) / WR, wc = (tid >> 6) % WR;
        const uint64_t mask = (uint64_t)im.lanemap[tid * 4] | ((uint64_t)im.lanemap[tid * 4 + 1] << 32);
        const int      base = (int)im.lanemap[tid * 4 + 2];
        for (int ti = 0; ti < TI; ++ti)
            for (int tj = 0; tj < TI; ++tj)
                for (int v = 0; v < 4; ++v)
                {
                    const int  slot = (ti * TI + tj) * 4 + v;
                    const int  ri = tile_row(wr + WR * ti, lr, v), ci = tile_col(wc + WR * tj, lc);
                    const bool nz = (mask >> slot) & 1;
                    const int  nzi = base + __builtin_popcountll(mask & ((1ull << slot) - 1ull));
                    if (nz) IN(tag + "fill sJv[nzi]", L::oJv + nzi, L::oJv, L::oJv + nnz);
                    IN(tag + "fill sJv[nz ? nzi : 0]", L::oJv + (nz ? nzi : 0), L::oJv, L::oJv + nnz);
                    if (ri >= 1 && ri < N && ci >= 1 && ci < N) mark(ri, ci);
                }
        if (wc == 0 && lc == 0) // column 0
            for (int ti = 0; ti < TI; ++ti)
                for (int v = 0; v < 4; ++v)
                {
                    const int ri = tile_row(wr + WR * ti, lr, v);
                    if (ri >= 1 && ri < N)
                    {
                        IN(tag + "column 0 sDYT[ri-1]", L::oDYdT + ri - 1, L::oDYdT, L::oCorr);
                        mark(ri, 0);
                    }
                }
        if (wc == JN % WR && lc == (N & 15)) // column N
            for (int ti = 0; ti < TI; ++ti)
                for (int v = 0; v < 4; ++v)
                {
                    const int ri = tile_row(wr + WR * ti, lr, v);
                    if (ri >= 1 && ri < N)
                    {
                        IN(tag + "column N sDYd[ri-1]", L::oDYdot + ri - 1, L::oDYdot, L::oDYdT);
                        IN(tag + "column N sDYT[ri-1]", L::oDYdT + ri - 1, L::oDYdT, L::oCorr);
                        IN(tag + "column N sCorr[ri-1]", L::oCorr + ri - 1, L::oCorr, L::oEv);
                        check(tile_col(wc + WR * (JN / WR), lc) == N, tag + "column N tile", ri, 0, 1);
                        mark(ri, N);
                    }
                }
        if (wr == 0 && lr == 0) // row 0
            for (int tj = 0; tj < TI; ++tj)
            {
                const int ci = tile_col(wc + WR * tj, lc);
                if (ci <= N)
                {
                    if (ci >= 1 && ci < N) IN(tag + "row 0 sBE[ci-1]", L::oBE + ci - 1, L::oBE, L::oDYdot);
                    if (MIX) IN(tag + "fill sRow0[ci]", L::oRow0 + ci, L::oRow0, L::oPan);
                    mark(0, ci);
                }
            }
    }
    for (int r = 0; r < N; ++r)
        for (int c = 0; c <= N; ++c)
            check(written[(size_t)r * (N + 1) + c] == 1, tag + "matrix entry written once", r * (N + 1) + c, 1, 2);

    // ---- tile LU: panel publish / factor steps, packed rows, multipliers ----
    const int panEnd = 16 * NPAD;
    IN(tag + "panel region end", L::oPan + panEnd, 0, total + 1);
    for (int fp32 = 0; fp32 <= (MIX ? 1 : 0); ++fp32)
    {
        // FP64: packed rows at sm[0, NU) in doubles.  FP32 (mixed): rows, multipliers and
        // reciprocals at sm[0, f32Words) in floats, panels reinterpret the panel doubles.
        const std::string t2   = tag + (fp32 ? "fp32 LU " : "fp64 LU ");
        const long        uEnd = fp32 ? (long)L::NU : (long)L::NU;               // elements of sU
        const long        lEnd = (long)L::NL, riEnd = N;                          // FP32 only
        const long        uLimit = fp32 ? 2L * L::oBE : (long)L::oY;              // what sU may reach
        check(uEnd <= uLimit, t2 + "packed rows below the live arrays", uEnd, 0, uLimit + 1);
        if (fp32) check(L::NU + lEnd + riEnd <= 2L * L::oBE, t2 + "factors below the tail", L::NU + lEnd + riEnd, 0, 2L * L::oBE + 1);
        // how often each packed-row / multiplier entry is stored (each must be stored exactly once)
        std::vector<int> uStores(uEnd, 0), lStores(lEnd, 0);
        auto             storeU = [&](long idx) { if (idx >= 0 && idx < uEnd) ++uStores[idx]; };
        auto             storeL = [&](long idx) { if (idx >= 0 && idx < lEnd) ++lStores[idx]; };
        for (int step = 0; step < X::NS; ++step)
        {
            const int p0 = 4 * step, Tp = p0 >> 4, c0 = p0 & 15, v0 = c0 >> 2, nq = (N - p0) < 4 ? (N - p0) : 4;
            for (int tid = 0; tid < NT; ++tid)
            {
                const int lr = (tid & 63) >> 4, lc = tid & 15, wr = (tid >> 6) / WR, wc = (tid >> 6) % WR;
                // publish<step>
                if (wc == Tp % WR && lc >= c0 && lc < c0 + 4)
                    for (int ti = 0; ti < TI; ++ti)
                        if (wr + WR * ti >= Tp)
                            for (int v = 0; v < 4; ++v)
                                IN(t2 + "publish P", X::pcol(step, X::trow(wr + WR * ti, lr, v), lc - c0), 0, panEnd);
                if (wr == Tp % WR)
                    for (int tj = 0; tj < TI; ++tj)
                        if (wc + WR * tj >= Tp) IN(t2 + "publish RT", X::prow(step, X::tcol(wc + WR * tj, lc), lr), 0, panEnd);
                (void)v0;
                // step<step>: diagonal block
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c)
                        IN(t2 + "diag P", X::pcol(step, p0 + r, c), 0, panEnd);
                if (tid == 0)
                    for (int q = 0; q < nq; ++q)
                    {
                        for (int c = q + 1; c < 4; ++c)
                            if (p0 + c <= N)
                            {
                                IN(t2 + "sU diag", X::uidx(p0 + q, p0 + c), 0, uEnd);
                                storeU(X::uidx(p0 + q, p0 + c));
                            }
                        if (fp32)
                        {
                            IN(t2 + "sRinv", p0 + q, 0, riEnd);
                            for (int c = 0; c < q; ++c)
                            {
                                IN(t2 + "sL diag", X::lrow(p0 + q) + p0 + c, 0, lEnd);
                                storeL(X::lrow(p0 + q) + p0 + c);
                            }
                        }
                    }
                const int HALF = NT / 2;
                if (tid < HALF)
                {
                    const int i = p0 + 4 + tid;
                    if (i < N)
                        for (int q = 0; q < 4; ++q)
                        {
                            IN(t2 + "row lane P", X::pcol(step, i, q), 0, panEnd);
                            if (fp32 && q < nq)
                            {
                                IN(t2 + "row lane sL", X::lrow(i) + p0 + q, 0, lEnd);
                                storeL(X::lrow(i) + p0 + q);
                            }
                        }
                }
                else
                {
                    const int j = p0 + 4 + (tid - HALF);
                    if (j <= N)
                        for (int q = 0; q < 4; ++q)
                        {
                            IN(t2 + "column lane RT", X::prow(step, j, q), 0, panEnd);
                            if (q < nq)
                            {
                                IN(t2 + "column lane sU", X::uidx(p0 + q, j), 0, uEnd);
                                storeU(X::uidx(p0 + q, j));
                            }
                        }
                }
                // MFMA operands: (i in range) ? -P(i, lr) : 0 and (j in range) ? RT(j, lr) : 0
                for (int ti = 0; ti < TI; ++ti)
                {
                    const int i = 16 * (wr + WR * ti) + X::arow(fp32, lc);
                    if (i >= p0 + 4 && i < N) IN(t2 + "A operand", X::pcol(step, i, lr), 0, panEnd);
                    IN(t2 + "A operand (select)", L::oPan * (fp32 ? 2 : 1) + X::pcol(step, i, lr), 0,
                       (long)total * (fp32 ? 2 : 1));
                }
                for (int tj = 0; tj < TI; ++tj)
                {
                    const int j = X::tcol(wc + WR * tj, lc);
                    if (j >= p0 + 4 && j <= N) IN(t2 + "B operand", X::prow(step, j, lr), 0, panEnd);
                    IN(t2 + "B operand (select)", L::oPan * (fp32 ? 2 : 1) + X::prow(step, j, lr), 0,
                       (long)total * (fp32 ? 2 : 1));
                }
            }
        }
        // every normalised-row entry (k, j), k < j <= N, and (mixed) every multiplier (i, c), c < i,
        // is stored by exactly one lane of exactly one block
        for (int kk = 0; kk < N; ++kk)
            for (int j = kk + 1; j <= N; ++j)
                check(uStores[X::uidx(kk, j)] == 1, t2 + "sU entry stored once", X::uidx(kk, j), 1, 2);
        if (fp32)
            for (int i = 1; i < N; ++i)
                for (int c = 0; c < i; ++c)
                    check(lStores[X::lrow(i) + c] == 1, t2 + "sL entry stored once", X::lrow(i) + c, 1, 2);
        // back substitution: right-hand side column, then the readlane chain's U loads
        for (int lane = 0; lane < 64; ++lane)
        {
            for (int a = 0; a < NA; ++a)
            {
                const int row = lane + 64 * a;
                IN(t2 + "rhs sU[uidx(row, N)]", row < N ? X::uidx(row, N) : 0, 0, uEnd);
            }
            for (int j0 = N - 1; j0 > 0; j0 -= 4)
                for (int s = 0; s < 4; ++s)
                    for (int a = 0; a < NA; ++a)
                    {
                        const int j = j0 - s, row = lane + 64 * a;
                        const bool ok = j > 0 && row < j;
                        IN(t2 + "back sU", ok ? X::uidx(row, j) : 0, 0, uEnd);
                    }
            if (fp32)
                for (int j0 = 0; j0 < N; j0 += 4)
                    for (int s = 0; s < 4; ++s)
                    {
                        const int j = j0 + s;
                        if (j < N) IN(t2 + "forward sRinv", j, 0, riEnd);
                        IN(t2 + "forward sRinv (select)", L::NU + L::NL + j, 0, 2L * total);
                        for (int a = 0; a < NA; ++a)
                        {
                            const int row = lane + 64 * a;
                            const bool ok = j < N && row > j && row < N;
                            IN(t2 + "forward sL", ok ? X::lrow(row) + j : 0, 0, lEnd);
                        }
                    }
        }
    }

    // ---- mixed refinement: residual in the panel buffers, row 0 dense, sparse rows ----
    if (MIX)
    {
        for (int ri = 0; ri < N; ++ri)
            IN(tag + "sRes[ri]", L::oPan + ri, L::oPan, total);
        for (int s = 0; s < D; ++s)
            for (int e = h16[im.oRowPtr16 + s]; e < h16[im.oRowPtr16 + s + 1]; ++e)
            {
                const int c = (int)(w32(im.oRowEnt + e) >> 16) + 1;
                IN(tag + "refine sX[c]", L::oX + c, L::oX, L::oX + N + 1);
            }
        // (row == 0 ? T : |sY[row - 1]|): the select may load sY[-1]
        IN(tag + "refine sY[row-1] (select)", L::oY - 1, 0, total);
    }
    // ---- the EOS lane's composition sums read sYb[0, zdim), zdim <= d ----
    for (int s = 0; s < D; ++s)
        IN(tag + "eos sYb", L::oYb + s, L::oYb, L::oX);
}
} // namespace

int main()
{
    run<14, false>();
    run<14, true>();
    run<86, false>();
    run<86, true>();
    run<87, false>();
    run<87, true>();
    std::printf("%d index checks, %d out of range\n", n_checks, n_fail);
    return n_fail ? 1 : 0;
}
