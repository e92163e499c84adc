// TEST INFRASTRUCTURE — C-ABI wrapper over the real SPH-EXA-nnet sources, compiled by
// oracle/Makefile.ref from the reference tree where it lies into oracle/_ref/libnnet_ref.so.
// Used only to pin the oracle restatement (bit-exact comparisons) and as the "reference" CPU
// baseline.  Nothing of the reference is copied: this file only calls its functions.
#include <cmath>
#include <omp.h>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nnet/nuclear_net.hpp"
#include "nnet_util/algorithm.hpp"
#include "nnet/parameterization/eos/ideal_gas.hpp"
#include "nnet/parameterization/net14/net14.hpp"
#include "nnet/parameterization/net86/net86.hpp"
#include "nnet/parameterization/net87/net87.hpp"
#ifdef NNET_REF_HELM
#include "nnet/parameterization/eos/helmholtz.hpp"
#include "nnet/parallel/implementation/parallel_nuclear_net_impl.hpp"
#endif

namespace
{
const nnet::ReactionList& list_for(int net)
{
    return net == 14 ? nnet::net14::reactionList : net == 86 ? nnet::net86::reactionList : nnet::net87::reactionList;
}
const nnet::ComputeReactionRatesFunctor<double>& rates_for(int net)
{
    if (net == 14) return nnet::net14::computeReactionRates;
    if (net == 86) return nnet::net86::computeReactionRates;
    return nnet::net87::computeReactionRates;
}
int dim_for(int net) { return net == 14 ? 14 : net == 86 ? 86 : 87; }

struct EosHolder
{
    nnet::eos::IdealGasFunctor<double> ideal;
#ifdef NNET_REF_HELM
    nnet::eos::HelmholtzFunctor<double> helm;
    EosHolder(double mu, const double* Z, int zdim)
        : ideal(mu)
        , helm(Z, zdim)
    {
    }
#else
    EosHolder(double mu, const double*, int)
        : ideal(mu)
    {
    }
#endif
    const nnet::EosFunctor<double>& get(int kind) const
    {
#ifdef NNET_REF_HELM
        if (kind == 1) return helm;
#endif
        return ideal;
    }
};

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// The reference's substep loop (nuclear_net.hpp:931-959) with its own prepareSystemSubstep,
// eigen::solve and finalizeSystemSubstep, plus the replay detection and Newton cap of the oracle
// (the reference itself would not return from a replaying particle).  Caller-owned scratch.
long guarded_particle(int net, const nnet::EosFunctor<double>& eos, double* Y, double* Tp, double rho, double drho_dt,
                      double dt_tot, double* dtp, long max_iter, double* stats, double* trace, int trace_cap,
                      double* rates, double* Mp, double* RHS, double* DY, double* Ybuf)
{
    const int             d = dim_for(net);
    nnet::PtrReactionList L(list_for(net));
    double&               T  = *Tp;
    double&               dt = *dtp;
    stats[0] = 0, stats[1] = 0, stats[2] = 0, stats[3] = 0;
    long len = 0;
    if (rho < nnet::constants::minRho || T < nnet::constants::minTemp) return 0;
    double elapsed = 0, T_buf = std::nan("");
    double att_dt = 0, att_T = 0;
    long   total = 0;
    for (int i = 1;; ++i)
    {
        const int it = i;
        if (i == 1) att_dt = dt, att_T = T_buf;
        double used = dt;
        if (dt_tot - elapsed < dt) used = dt_tot - elapsed;
        const double el0 = elapsed;
        try
        {
            nnet::prepareSystemSubstep(d, Mp, RHS, rates, L, rates_for(net), eos, Y, T, Ybuf, T_buf, rho, drho_dt,
                                       dt_tot, elapsed, dt, i);
        }
        catch (...)
        {
            stats[0] = 3;
            return len;
        }
        eigen::solve(Mp, RHS, DY, d + 1, (double)nnet::constants::epsilonSystem);
        stats[2] += 1;
        ++total;
        const bool done = nnet::finalizeSystemSubstep(d, Y, T, Ybuf, T_buf, DY, dt_tot, elapsed, dt, i);
        const bool conv = (i == 0) && (elapsed != el0 || done);
        if (len < trace_cap && trace)
        {
            double* p = trace + 6 * len;
            p[0] = it, p[1] = used, p[2] = el0, p[3] = dt, p[4] = T_buf, p[5] = conv ? 1 : 0;
        }
        ++len;
        if (conv) stats[1] += 1;
        if (done) return len;
        if (i == 0 && !conv)
        {
            stats[3] += 1;
            if (same_bits(dt, att_dt) && same_bits(T_buf, att_T))
            {
                stats[0] = 2;
                return len;
            }
        }
        if (max_iter > 0 && total >= max_iter)
        {
            stats[0] = 4;
            return len;
        }
    }
}
} // namespace

extern "C"
{
    int ref_has_helmholtz()
    {
#ifdef NNET_REF_HELM
        return 1;
#else
        return 0;
#endif
    }

    int ref_num_reactions(int net) { return (int)list_for(net).size(); }

    // Flattened reaction list through the public operator[] (nuclear_net.hpp:199-209).
    int ref_reaction_list(int net, int* reactant_begin, int* product_begin, int* ids, int* nums)
    {
        const nnet::ReactionList& L = list_for(net);
        int                       k = 0;
        for (size_t r = 0; r < L.size(); ++r)
        {
            auto R            = L[(int)r];
            reactant_begin[r] = k;
            for (auto [id, n] : R.reactants)
                ids[k] = id, nums[k] = n, ++k;
            product_begin[r] = k;
            for (auto [id, n] : R.products)
                ids[k] = id, nums[k] = n, ++k;
        }
        reactant_begin[L.size()] = k;
        return k;
    }

    void ref_rates(int net, const double* Y, double T, double rho, double* BE, double* rates, double* drates)
    {
        nnet::eos_struct<double> eos{};
        eos.dudYe = 0;
        rates_for(net)(Y, T, rho, eos, BE, rates, drates);
    }

    void ref_solve(double* M, double* RHS, double* X, int n) { eigen::solve(M, RHS, X, n, (double)1e-40); }

    void ref_eos(int kind, double mu, const double* Z, int zdim, const double* Y, double T, double rho, double* out)
    {
        EosHolder h(mu, Z, zdim);
        auto      r = h.get(kind)(Y, T, rho);
        const double v[13] = {r.cv, r.dpdT, r.p, r.cp, r.c, r.u, r.dse, r.dpe, r.dsp, r.cv_gaz, r.cp_gaz, r.c_gaz, r.dudYe};
        std::memcpy(out, v, sizeof(v));
    }

    int ref_prepare(int net, int eos_kind, double mu, const double* Z, int zdim, const double* Y, double T,
                    const double* Yg, double Tg, double rho, double drho_dt, double dt, double* Mp, double* RHS)
    {
        EosHolder                h(mu, Z, zdim);
        auto                     eos = h.get(eos_kind)(Yg, Tg, rho);
        std::vector<double>      rates(list_for(net).size());
        nnet::PtrReactionList    L(list_for(net));
        try
        {
            nnet::prepareSystemFromGuess(dim_for(net), Mp, RHS, rates.data(), L, rates_for(net), Y, T, Yg, Tg, rho,
                                         drho_dt, eos, dt);
        }
        catch (...)
        {
            return 1;
        }
        return 0;
    }

    // The reference's substep loop body (prepareSystemSubstep, eigen::solve,
    // finalizeSystemSubstep; nuclear_net.hpp:944-958) driven step by step, with the same trace
    // and the same replay detection as the oracle, so stalling inputs can be compared too.
    long ref_particle(int net, int eos_kind, double mu, const double* Z, int zdim, double* Y, double* Tp, double rho,
                      double drho_dt, double dt_tot, double* dtp, long max_iter, double* stats, double* trace,
                      int trace_cap)
    {
        const int           d = dim_for(net);
        EosHolder           h(mu, Z, zdim);
        std::vector<double> rates(list_for(net).size()), Mp((d + 1) * (d + 1)), RHS(d + 1), DY(d + 1), Ybuf(d);
        return guarded_particle(net, h.get(eos_kind), Y, Tp, rho, drho_dt, dt_tot, dtp, max_iter, stats, trace,
                                trace_cap, rates.data(), Mp.data(), RHS.data(), DY.data(), Ybuf.data());
    }

    // The reference's OpenMP particle loop (parallel_nuclear_net_impl.cpp:63-103: skip test,
    // schedule(dynamic, util::dynamicBatchSize), per-thread buffers) over guarded_particle, i.e.
    // the reference's compiled substep solver with the replay detection and Newton cap that let it
    // return on every input.  status: 0 ok, 1 skipped, 2 stalled, 3 zero dt, 4 capped.
    void ref_batch_guarded(int net, int eos_kind, double mu, const double* Z, int zdim, size_t n, const double* rho,
                           double** Y, double* temp, double* dt, double hydro_dt, long max_iter, int nthreads,
                           uint8_t* status, uint32_t* newton)
    {
        const int d  = dim_for(net);
        const int nR = (int)list_for(net).size();
        EosHolder h(mu, Z, zdim);
        if (nthreads > 0) omp_set_num_threads(nthreads);
        int num_threads = 1;
#pragma omp parallel
#pragma omp master
        num_threads = omp_get_num_threads();
        const int chunk = util::dynamicBatchSize(n, num_threads);
#pragma omp parallel
        {
            std::vector<double> rates(nR), Mp((d + 1) * (d + 1)), RHS(d + 1), DY(d + 1), Ybuf(d), Yl(d);
            double              stats[4];
#pragma omp for schedule(dynamic, chunk)
            for (size_t i = 0; i < n; ++i)
            {
                uint8_t  st = 1;
                uint32_t nw = 0;
                if (rho[i] > nnet::constants::minRho && temp[i] > nnet::constants::minTemp)
                {
                    for (int j = 0; j < d; ++j)
                        Yl[j] = Y[j][i];
                    guarded_particle(net, h.get(eos_kind), Yl.data(), &temp[i], rho[i], 0., hydro_dt, &dt[i],
                                     max_iter, stats, nullptr, 0, rates.data(), Mp.data(), RHS.data(), DY.data(),
                                     Ybuf.data());
                    for (int j = 0; j < d; ++j)
                        Y[j][i] = Yl[j];
                    st = (uint8_t)stats[0];
                    nw = (uint32_t)stats[2];
                }
                if (status) status[i] = st;
                if (newton) newton[i] = nw;
            }
        }
    }

    // The reference's own solveSystemSubstep (nuclear_net.hpp:968-979).  Never returns for
    // inputs on which the controller replays forever: call it only on particles the oracle
    // has shown to finish.
    void ref_solve_substep(int net, int eos_kind, double mu, const double* Z, int zdim, double* Y, double* T, double rho,
                           double drho_dt, double dt_tot, double* dt)
    {
        EosHolder h(mu, Z, zdim);
        nnet::solveSystemSubstep(dim_for(net), list_for(net), rates_for(net), h.get(eos_kind), Y, *T, rho, drho_dt,
                                 dt_tot, *dt);
    }

#ifdef NNET_REF_HELM
    // The reference's OpenMP batch path (parallel_nuclear_net_impl.cpp:63-103).  Same caveat
    // about replaying particles as ref_solve_substep.
    void ref_batch(int net, int eos_kind, double mu, const double* Z, int zdim, size_t n, double* rho, double* rho_m1,
                   double** Y, double* temp, double* dt, double hydro_dt, double previous_dt, int use_drhodt)
    {
        EosHolder h(mu, Z, zdim);
        nnet::parallel::computeNuclearReactionsImpl(n, dim_for(net), rho, rho_m1, Y, temp, dt, hydro_dt, previous_dt,
                                                    list_for(net), rates_for(net), h.get(eos_kind), use_drhodt,
                                                    (void*)nullptr);
    }

    // The reference's OpenMP post-step EOS refresh (parallel_nuclear_net_impl.cpp:116-141).
    void ref_helm_eos_batch(size_t n, int dim, const double* Z, const double* temp, const double* rho, double** Y,
                            double* u, double* cv, double* p, double* c, double* dpdT, int nthreads)
    {
        if (nthreads > 0) omp_set_num_threads(nthreads);
        nnet::parallel::computeHelmholtzImpl(n, dim, Z, temp, rho, Y, u, cv, p, c, dpdT);
    }
#endif
}
