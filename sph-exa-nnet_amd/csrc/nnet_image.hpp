// Host-side construction of a network's assembly image: the reaction list flattened into the
// compact tables that every solve workgroup stages into LDS (reaction CSR, per-species signed
// count lists, per-Jacobian-non-zero contribution lists, per-row non-zero lists, the order in
// which lanes take the non-zeros) and the per-lane map of the register slots that hold Jacobian
// non-zeros in the tile layout of the LU.  Also the library's built-in reaction lists.  Plain
// C++ (no HIP), so host-side checks can build the same images the library uploads.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "nnet_layout.hpp"
#include "nnet_tables.h"

namespace nnet_amd
{
// The library's reaction lists (net14.hpp:69-134, net86.hpp:95-192, net87.hpp:88-98).
struct HostList
{
    std::vector<int> rb{0}, pb, ids, nums;
    void push(std::vector<std::pair<int, int>> r, std::vector<std::pair<int, int>> pr)
    {
        for (auto& x : r)
            ids.push_back(x.first), nums.push_back(x.second);
        for (auto& x : pr)
            ids.push_back(x.first), nums.push_back(x.second);
        pb.push_back(rb.back() + (int)r.size());
        rb.push_back(pb.back() + (int)pr.size());
    }
};

inline HostList builtin_list(int kind)
{
    HostList L;
    if (kind == 14)
    {
        L.push({{0, 3}}, {{1, 1}});
        L.push({{0, 1}, {1, 1}}, {{2, 1}});
        L.push({{0, 1}, {2, 1}}, {{3, 1}});
        for (int s = 3; s < 13; ++s)
            L.push({{0, 1}, {s, 1}}, {{s + 1, 1}});
        L.push({{1, 2}}, {{3, 1}, {0, 1}});
        L.push({{1, 1}, {2, 1}}, {{4, 1}, {0, 1}});
        L.push({{2, 2}}, {{5, 1}, {0, 1}});
        L.push({{1, 1}}, {{0, 3}});
        L.push({{2, 1}}, {{0, 1}, {1, 1}});
        L.push({{3, 1}}, {{0, 1}, {2, 1}});
        for (int s = 4; s < 14; ++s)
            L.push({{s, 1}}, {{0, 1}, {s - 1, 1}});
        return L;
    }
    const int* mr = NNET86_MAIN_REACTANT;
    const int* sr = NNET86_SECONDARY_REACTANT;
    const int* mp = NNET86_MAIN_PRODUCT;
    auto       other = [](int dZ, int dA) -> int
    {
        if (dZ == 0 && dA == 0) return -1;
        if (dA == 1 && dZ == 0) return 1; // neutron
        if (dA == 1 && dZ == 1) return 0; // proton
        if (dA == 4 && dZ == 2) return 2; // alpha
        return -2;
    };
    L.push({{mr[0], 2}}, {{mp[0], 1}, {2, 1}});
    L.push({{mr[1], 1}, {sr[1], 1}}, {{mp[1], 1}, {2, 1}});
    L.push({{mr[2], 2}}, {{mp[2], 1}, {2, 1}});
    L.push({{mr[4], 3}}, {{mp[4], 1}});
    for (int i = 5; i < 157; ++i)
    {
        const int o = other((int)(NNET86_Z[mr[i]] + NNET86_Z[sr[i]] - NNET86_Z[mp[i]]),
                            (int)(NNET86_A[mr[i]] + NNET86_A[sr[i]] - NNET86_A[mp[i]]));
        if (o == -2) return HostList{};
        if (o < 0) L.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}});
        else
            L.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}, {o, 1}});
    }
    L.push({{mp[4], 1}}, {{mr[4], 3}});
    for (int i = 5; i < 157; ++i)
    {
        const int o = other((int)((NNET86_Z[mr[i]] + NNET86_Z[sr[i]]) - NNET86_Z[mp[i]]),
                            (int)((NNET86_A[mr[i]] + NNET86_A[sr[i]]) - NNET86_A[mp[i]]));
        if (o == -2) return HostList{};
        if (o < 0) L.push({{mp[i], 1}}, {{mr[i], 1}, {sr[i], 1}});
        else
            L.push({{mp[i], 1}, {o, 1}}, {{mr[i], 1}, {sr[i], 1}});
    }
    if (kind == 87)
    {
        L.push({{0, 1}, {86, 1}}, {{1, 1}});
        L.push({{1, 1}, {86, 1}}, {{0, 1}});
    }
    return L;
}


// Rate-parameterization tables of a network kind as the device kernels read them: constant binding
// energies and Z^(5/3) per species, rate-fit coefficients, Q values, partition-function rows, the
// choose rows of each fitted rate's reverse (ca / cb, photodisintegration flag), per rate-slot the
// screening species (a, b, c, optional e), forward/reverse swap and screened-reverse flags, and the
// reaction-list positions the slots are pushed to (net14.cpp:40-660, net86.cpp:41-686).
struct RateTables
{
    std::vector<double> BE0, Z53, fit, q, choose;
    std::vector<int>    ca, cb, photo, scr, swp, scl, pe, pl;
    int                 nfit = 0, nE = 0, e0 = 0;
    SpecialIdx          sp{};
};

inline RateTables build_rate_tables(int kind)
{
    RateTables t;
    const int  Dr = kind == 14 ? 14 : 86;
    std::vector<double> Zr(Dr);
    auto& BE0 = t.BE0;
    auto& Z53 = t.Z53;
    BE0.resize(Dr);
    Z53.resize(Dr);
    for (int s = 0; s < Dr; ++s)
    {
        BE0[s] = (kind == 14 ? NNET14_BE_MEV[s] : NNET86_BE_MEV[s]) * NNET_MEV_TO_ERG;
        Zr[s]  = kind == 14 ? NNET14_Z[s] : NNET86_Z[s];
        Z53[s] = std::pow(Zr[s], 5. / 3.);
    }
    const int nfit = kind == 14 ? 10 : 150, nE = kind == 14 ? 16 : 157;
    t.nfit = nfit, t.nE = nE, t.e0 = kind == 14 ? 3 : 7;
    auto &fit = t.fit, &q = t.q, &choose = t.choose;
    auto &ca = t.ca, &cb = t.cb, &photo = t.photo, &scr = t.scr, &swp = t.swp, &scl = t.scl, &pe = t.pe, &pl = t.pl;
    fit.resize(nfit * 8), q.resize(nfit), ca.resize(nfit), cb.resize(nfit), photo.resize(nfit);
    scr.resize(4 * nE), swp.assign(nE, 0), scl.assign(nE, 0), pe.assign(nE, -1), pl.assign(nE, -1);
    if (kind == 14)
    {
        std::memcpy(fit.data(), NNET14_FIT, sizeof(NNET14_FIT));
        std::memcpy(q.data(), NNET14_Q, sizeof(NNET14_Q));
        choose.assign(&NNET14_CHOOSE[0][0], &NNET14_CHOOSE[0][0] + 11 * 24);
        for (int f = 0; f < nfit; ++f)
            ca[f] = f, cb[f] = f + 1, photo[f] = 1;
        for (int i = 0; i < 14; ++i)
            scr[4 * i] = i, scr[4 * i + 1] = 0, scr[4 * i + 2] = i + 1, scr[4 * i + 3] = -1;
        scr[3] = 0;
        int o13[4] = {1, 1, 3, -1}, o14[4] = {1, 2, 4, -1}, o15[4] = {2, 2, 5, -1};
        std::memcpy(&scr[4 * 13], o13, sizeof(o13));
        std::memcpy(&scr[4 * 14], o14, sizeof(o14));
        std::memcpy(&scr[4 * 15], o15, sizeof(o15));
        for (int i = 0; i < 16; ++i)
            pe[i] = i;
        for (int i = 0; i < 13; ++i)
            pl[i] = 16 + i;
    }
    else
    {
        std::memcpy(fit.data(), NNET86_FIT, sizeof(NNET86_FIT));
        std::memcpy(q.data(), NNET86_Q, sizeof(NNET86_Q));
        choose.assign(&NNET86_CHOOSE[0][0], &NNET86_CHOOSE[0][0] + 81 * 24);
        for (int f = 0; f < nfit; ++f)
        {
            const int i = f + 7;
            ca[f]       = NNET86_MAIN_REACTANT[i] - 5;
            cb[f]       = NNET86_MAIN_PRODUCT[i] - 5;
            photo[f]    = i < 137 ? 1 : 0;
        }
        for (int i = 0; i < 157; ++i)
        {
            scr[4 * i]     = NNET86_MAIN_REACTANT[i];
            scr[4 * i + 1] = NNET86_SECONDARY_REACTANT[i];
            scr[4 * i + 2] = NNET86_MAIN_PRODUCT[i];
            scr[4 * i + 3] = i == 4 ? NNET86_MAIN_REACTANT[4] : -1;
            swp[i]         = (i >= 137 && (i < 154 || i == 156)) ? 1 : 0;
            scl[i]         = i >= 137 ? 1 : 0;
        }
        int idx = 0;
        for (int i = 0; i < 157; ++i)
            if (i != 3) pe[i] = idx++;
        for (int i = 4; i < 157; ++i)
            pl[i] = idx++;
    }
    t.sp = kind == 14 ? SpecialIdx{0, 0, 13, 14, 15, 1, 1, 2, 2} : SpecialIdx{4, 4, 0, 1, 2, 5, 5, 6, 6};
    return t;

}

// Word offsets (halfword offsets for the *16 arrays) of the image's sub-arrays, as NetDev keeps them.
struct AsmImage
{
    std::vector<uint32_t> img;     // the image, 32-bit words
    std::vector<uint32_t> lanemap; // [lane][4]: register-slot mask (lo, hi), first non-zero id, 0
    int                   nnz = 0;
    int oRx = 0, oEnt16 = 0, oSpPtr16 = 0, oSpEnt = 0, oNzPtr16 = 0, oNzEnt = 0, oRowPtr16 = 0, oRowEnt = 0, oJord16 = 0,
        oMaPerm16 = 0;
};

// Builds the image of a validated reaction list (reaction r: reactants [rb[r], pb[r]) and
// products [pb[r], rb[r+1]) of (ids, nums)) for rate parameterization `kind` with d species.
// Returns an empty string, or the reason the list does not fit the kernel's capacities.
inline std::string build_asm_image(int kind, int d, int nR, const int* rb, const int* pb, const int* ids,
                                   const int* nums, AsmImage& out)
{
    // derivativesFromReactions: per species, the ordered (reaction, signed count) updates
    std::vector<std::vector<std::pair<int, int>>> per(d);
    for (int r = 0; r < nR; ++r)
    {
        for (int qq = rb[r]; qq < pb[r]; ++qq)
            per[ids[qq]].push_back({r, -nums[qq]});
        for (int qq = pb[r]; qq < rb[r + 1]; ++qq)
            per[ids[qq]].push_back({r, nums[qq]});
    }
    // firstOrderDYFromReactions: per matrix entry (a, b), the ordered (reaction, reactant
    // position of b, signed count) updates, in the reference's loop order
    struct Contrib
    {
        int r, q, c;
    };
    std::vector<std::vector<Contrib>> ent((size_t)d * d);
    for (int r = 0; r < nR; ++r)
        for (int qb = rb[r]; qb < pb[r]; ++qb)
        {
            const int b = ids[qb];
            for (int o = rb[r]; o < pb[r]; ++o)
                ent[(size_t)ids[o] * d + b].push_back({r, qb - rb[r], -nums[o]});
            for (int o = pb[r]; o < rb[r + 1]; ++o)
                ent[(size_t)ids[o] * d + b].push_back({r, qb - rb[r], nums[o]});
        }
    // Non-zeros are numbered in the order of the register slots that hold them (lane-major, then
    // slot), so each lane's entries are consecutive: a lane finds its k-th entry at its base + k.
    const int NT = kind == 14 ? Cfg<14>::NT : Cfg<86>::NT, TT = kind == 14 ? Cfg<14>::TT : Cfg<86>::TT,
              WR = kind == 14 ? Cfg<14>::WR : Cfg<86>::WR, TI = TT / WR;
    std::vector<int>      nz_of((size_t)d * d, -1), nz_row, nz_col;
    std::vector<uint32_t>& lanemap = out.lanemap; // per lane: slot mask (lo, hi), base, unused
    lanemap.assign((size_t)NT * 4, 0u);
    for (int t = 0; t < NT; ++t)
    {
        const int l = t & 63, wr = (t >> 6) / WR, wc = (t >> 6) % WR;
        lanemap[(size_t)t * 4 + 2] = (uint32_t)nz_row.size();
        for (int ti = 0; ti < TI; ++ti)
            for (int tj = 0; tj < TI; ++tj)
                for (int v = 0; v < 4; ++v)
                {
                    const int ri = tile_row(wr + WR * ti, l >> 4, v), ci = tile_col(wc + WR * tj, l & 15);
                    if (ri < 1 || ri > d || ci < 1 || ci > d || ent[(size_t)(ri - 1) * d + (ci - 1)].empty()) continue;
                    const int slot = (ti * TI + tj) * 4 + v;
                    lanemap[(size_t)t * 4 + (slot >> 5)] |= 1u << (slot & 31);
                    nz_of[(size_t)(ri - 1) * d + (ci - 1)] = (int)nz_row.size();
                    nz_row.push_back(ri - 1);
                    nz_col.push_back(ci - 1);
                }
    }
    const int nnz = (int)nz_row.size();
    out.nnz       = nnz;

    std::vector<uint32_t>& img = out.img;
    img.clear();
    // rx
    out.oRx = (int)img.size();
    for (int r = 0; r < nR; ++r)
        img.push_back((uint32_t)rb[r] | ((uint32_t)pb[r] << 16));
    img.push_back((uint32_t)rb[nR]);
    // entries (halfwords)
    out.oEnt16 = (int)img.size() * 2;
    std::vector<uint16_t> h;
    for (int q = 0; q < rb[nR]; ++q)
        h.push_back((uint16_t)(ids[q] | (nums[q] << 8)));
    auto push_half = [&](std::vector<uint16_t>& v)
    {
        if (v.size() & 1) v.push_back(0);
        for (size_t i = 0; i < v.size(); i += 2)
            img.push_back((uint32_t)v[i] | ((uint32_t)v[i + 1] << 16));
    };
    push_half(h);
    // species lists
    out.oSpPtr16 = (int)img.size() * 2;
    h.clear();
    {
        int acc = 0;
        for (int s = 0; s < d; ++s)
        {
            h.push_back((uint16_t)acc);
            acc += (int)per[s].size();
        }
        h.push_back((uint16_t)acc);
    }
    push_half(h);
    out.oSpEnt = (int)img.size();
    for (int s = 0; s < d; ++s)
        for (auto& x : per[s])
            img.push_back((uint32_t)x.first | (((uint32_t)(uint8_t)(int8_t)x.second) << 16));
    // non-zeros
    out.oNzPtr16 = (int)img.size() * 2;
    h.clear();
    {
        int acc = 0;
        for (int e = 0; e < nnz; ++e)
        {
            h.push_back((uint16_t)acc);
            acc += (int)ent[(size_t)nz_row[e] * d + nz_col[e]].size();
        }
        h.push_back((uint16_t)acc);
    }
    push_half(h);
    out.oNzEnt = (int)img.size();
    for (int e = 0; e < nnz; ++e)
        for (auto& x : ent[(size_t)nz_row[e] * d + nz_col[e]])
            img.push_back((uint32_t)x.r | ((uint32_t)x.q << 12) | (((uint32_t)(uint8_t)(int8_t)x.c) << 16));
    // per row, its non-zeros in ascending column order: (id, column)
    out.oRowPtr16 = (int)img.size() * 2;
    h.clear();
    std::vector<uint32_t> rowent;
    for (int a = 0; a < d; ++a)
    {
        h.push_back((uint16_t)rowent.size());
        for (int b = 0; b < d; ++b)
            if (nz_of[(size_t)a * d + b] >= 0) rowent.push_back((uint32_t)nz_of[(size_t)a * d + b] | ((uint32_t)b << 16));
    }
    h.push_back((uint16_t)rowent.size());
    push_half(h);
    out.oRowEnt = (int)img.size();
    img.insert(img.end(), rowent.begin(), rowent.end());
    // compute order of the non-zeros: descending cost (8-wide load groups + serial remainder)
    out.oJord16 = (int)img.size() * 2;
    {
        std::vector<int> cost(nnz);
        for (int e = 0; e < nnz; ++e)
        {
            const int len = (int)ent[(size_t)nz_row[e] * d + nz_col[e]].size();
            cost[e]       = len / 8 + len % 8;
        }
        h.clear();
        for (int e = 0; e < nnz; ++e)
            h.push_back((uint16_t)e);
        std::stable_sort(h.begin(), h.end(), [&](uint16_t a, uint16_t b) { return cost[a] > cost[b]; });
        push_half(h);
    }
    // mass-action order: slot s of the per-reaction factor phase (lane s % NT, pass s / NT)
    // computes reaction perm[s].  A wavefront runs as many passes as its busiest lane, so the
    // wavefronts with the most passes take the cheapest reactions (fewest reactant entries, no
    // count above one: one load chain, no division) in all their passes; with a single pass
    // everywhere the order is the identity.
    out.oMaPerm16 = (int)img.size() * 2;
    {
        const int        NTk = kind == 14 ? Cfg<14>::NT : Cfg<86>::NT;
        std::vector<int> perm(nR);
        for (int r = 0; r < nR; ++r)
            perm[r] = r;
        if (nR > NTk)
        {
            std::vector<int> rcost(nR), order(nR), slots(nR), passes((NTk + 63) / 64, 0);
            for (int r = 0; r < nR; ++r)
            {
                int c2 = 0;
                for (int k = rb[r]; k < pb[r]; ++k)
                    c2 += nums[k] > 1;
                rcost[r] = (pb[r] - rb[r]) + 2 * c2;
                order[r] = r;
            }
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rcost[a] < rcost[b]; });
            for (int s2 = 0; s2 < nR; ++s2)
            {
                passes[(s2 % NTk) / 64] = s2 / NTk + 1;
                slots[s2]               = s2;
            }
            std::stable_sort(slots.begin(), slots.end(),
                             [&](int a, int b) { return passes[(a % NTk) / 64] > passes[(b % NTk) / 64]; });
            for (int i = 0; i < nR; ++i)
                perm[slots[i]] = order[i];
        }
        h.clear();
        for (int r = 0; r < nR; ++r)
            h.push_back((uint16_t)perm[r]);
        push_half(h);
    }

    const int img_cap = kind == 14 ? Cfg<14>::IMG : Cfg<86>::IMG;
    const int nnz_cap = kind == 14 ? Cfg<14>::NNZ : Cfg<86>::NNZ;
    if ((int)img.size() > img_cap || nnz >= nnz_cap || rb[nR] > 65535)
        return "reaction list too large for the kernel's LDS image";
    return std::string();
}
} // namespace nnet_amd
