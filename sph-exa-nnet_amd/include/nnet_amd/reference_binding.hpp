/* Binding of the reference's own functor objects to the MI355X solve.
 *
 * Include after the SPH-EXA-nnet headers (this file includes the ones it needs, so the
 * reference's include/ directory must be on the include path).  It specializes the functor traits
 * of nnet_amd/parallel_nuclear_net.hpp for every type derived from
 *   nnet::ComputeReactionRatesFunctor<double>  (nuclear_net.hpp:483-503)
 *   nnet::EosFunctor<double>                   (nuclear_net.hpp:463-479)
 * by their dynamic type, the way a maintainer replaces the reference's GPU branch
 * (parallel_nuclear_net.hpp:95-142, which copies these objects to the device):
 *   nnet::net14/net86/net87::ComputeReactionRatesFunctor<double>  -> NNET_RATES_NET14/86/87
 *   nnet::eos::IdealGasFunctor<double>(mu)                       -> NNET_EOS_IDEAL, mu
 *   nnet::eos::HelmholtzFunctor<double>(Z, dimension)            -> NNET_EOS_HELMHOLTZ, Z, dimension
 * Any other dynamic type throws nnet_amd::Error(NNET_E_ARG): there is no device implementation of
 * it and no CPU fallback.  With it, the reference's call
 *   nnet::parallel::computeNuclearReactions<true>(n, first, last, hydro_dt, previous_dt,
 *       reactionList, construct_rates_BE, eos, use_drhodt);
 * becomes, with the same arguments (the reference's nnet::ReactionList included),
 *   nnet_amd::parallel::computeNuclearReactions<true>(n, first, last, hydro_dt, previous_dt,
 *       reactionList, construct_rates_BE, eos, use_drhodt);
 *
 * IdealGasFunctor::mu and HelmholtzFunctor::Z / dimension are private members of the reference's
 * classes and the classes offer no accessors.  They are read through pointers to member formed in
 * explicit instantiations, where the language does not apply access checking ([temp.spec]), so
 * the reference headers need no edit.
 */
#pragma once

#include <type_traits>

#include "nnet/nuclear_net.hpp"
#include "nnet/parameterization/eos/helmholtz.hpp"
#include "nnet/parameterization/eos/ideal_gas.hpp"
#include "nnet/parameterization/net14/net14.hpp"
#include "nnet/parameterization/net86/net86.hpp"
#include "nnet/parameterization/net87/net87.hpp"

#include "parallel_nuclear_net.hpp"

namespace nnet_amd
{
namespace ref_member
{
// member(Tag{}) returns the pointer to member named in the explicit instantiation below
template<class Tag, typename Tag::type M>
struct Grant
{
    friend typename Tag::type member(Tag) { return M; }
};
struct IdealMu
{
    using type = double nnet::eos::IdealGasFunctor<double>::*;
    friend type member(IdealMu);
};
struct HelmZ
{
    using type = const double* nnet::eos::HelmholtzFunctor<double>::*;
    friend type member(HelmZ);
};
struct HelmDim
{
    using type = int nnet::eos::HelmholtzFunctor<double>::*;
    friend type member(HelmDim);
};
template struct Grant<IdealMu, &nnet::eos::IdealGasFunctor<double>::mu>;
template struct Grant<HelmZ, &nnet::eos::HelmholtzFunctor<double>::Z>;
template struct Grant<HelmDim, &nnet::eos::HelmholtzFunctor<double>::dimension>;
} // namespace ref_member

/* The rate parameterization of a reference rate functor (its dynamic type). */
inline int reference_rates_kind(const nnet::ComputeReactionRatesFunctor<double>& f)
{
    if (dynamic_cast<const nnet::net87::ComputeReactionRatesFunctor<double>*>(&f)) return NNET_RATES_NET87;
    if (dynamic_cast<const nnet::net86::ComputeReactionRatesFunctor<double>*>(&f)) return NNET_RATES_NET86;
    if (dynamic_cast<const nnet::net14::ComputeReactionRatesFunctor<double>*>(&f)) return NNET_RATES_NET14;
    throw Error(NNET_E_ARG, "rate functor of a type without a device implementation (net14/net86/net87 only)");
}

/* The EOS description of a reference EOS functor (its dynamic type). */
inline void reference_eos_describe(const nnet::EosFunctor<double>& e, nnet_batch& b)
{
    if (auto* g = dynamic_cast<const nnet::eos::IdealGasFunctor<double>*>(&e))
    {
        b.eos = NNET_EOS_IDEAL;
        b.mu  = g->*member(ref_member::IdealMu{});
        return;
    }
    if (auto* h = dynamic_cast<const nnet::eos::HelmholtzFunctor<double>*>(&e))
    {
        b.eos   = NNET_EOS_HELMHOLTZ;
        b.Z     = h->*member(ref_member::HelmZ{});
        b.z_dim = h->*member(ref_member::HelmDim{});
        return;
    }
    throw Error(NNET_E_ARG, "EOS functor of a type without a device implementation (ideal gas, Helmholtz only)");
}

template<class T>
struct RatesKind<T, std::enable_if_t<std::is_base_of_v<nnet::ComputeReactionRatesFunctor<double>, T>>>
{
    static int kind(const T& f) { return reference_rates_kind(f); }
};

template<class T>
struct EosDesc<T, std::enable_if_t<std::is_base_of_v<nnet::EosFunctor<double>, T>>>
{
    static void describe(const T& e, nnet_batch& b) { reference_eos_describe(e, b); }
};
} // namespace nnet_amd
