/* C++ host interface of the MI355X nuclear-network solve, mirroring the reference's
 *   nnet::parallel::computeNuclearReactions(Data& n, size_t firstIndex, size_t lastIndex,
 *       Float hydro_dt, Float previous_dt, const ReactionList& reactions,
 *       const ComputeReactionRatesFunctor<Float>& construct_rates_BE,
 *       const EosFunctor<Float>& eos, bool use_drhodt, nseFunction jumpToNse)
 *                                       (include/nnet/parallel/parallel_nuclear_net.hpp:86-172)
 * with the same argument meaning, on top of the C-ABI of include/nnet_mi355x.h.
 *
 * What a caller keeps from the reference:
 *   - the reaction list (Reaction / ReactionList with pushBack, operator[], size(); any list type
 *     with that interface, including the reference's nnet::ReactionList, is accepted),
 *   - the rate-functor choice (net14 / net86 / net87) and the EOS functor (ideal gas with mu,
 *     Helmholtz with the species charges),
 *   - the Data requirements: fields rho, rho_m1, temp, dt and Y (a vector of per-species arrays)
 *     of n.devData, each exposing .data() (a raw pointer or a thrust-style device pointer with
 *     .get()), n.numSpecies, and for use_drhodt an allocated rho_m1.
 * All particle arrays stay in device memory; the call is stream-ordered on the default stream
 * and synchronises before it returns, like the reference's GPU branch.
 *
 * Functor objects are mapped to the device parameterization by two traits, RatesKind<T> and
 * EosDesc<T> (below).  They are specialized here for this header's functor types;
 * nnet_amd/reference_binding.hpp specializes them for the reference's own polymorphic objects
 * (nnet::net14/86/87::ComputeReactionRatesFunctor<double>, nnet::eos::IdealGasFunctor<double>,
 * nnet::eos::HelmholtzFunctor<double>, seen through their nnet:: base classes), and a caller can
 * specialize them for functor types of its own.
 *
 * jumpToNse is accepted for signature compatibility and ignored, as in the reference
 * (parallel_nuclear_net_impl.cpp:63-103 never calls it for the solve).
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../../include/nnet_mi355x.h"

namespace nnet_amd
{
/* ---- reaction list (nuclear_net.hpp:111-213) ---- */
struct Reaction
{
    struct ReactantProduct
    {
        int speciesId, numConsumed = 1;
    };
    std::vector<ReactantProduct> reactants, products;
};

class ReactionList
{
public:
    struct Span
    {
        const Reaction::ReactantProduct* b;
        const Reaction::ReactantProduct* e;
        const Reaction::ReactantProduct* begin() const { return b; }
        const Reaction::ReactantProduct* end() const { return e; }
        size_t                           size() const { return (size_t)(e - b); }
    };
    struct Ref
    {
        Span reactants, products;
    };

    ReactionList() = default;
    explicit ReactionList(const std::vector<Reaction>& reactions)
    {
        for (auto& r : reactions)
            pushBack(r);
    }
    void pushBack(const Reaction& r)
    {
        entries_.insert(entries_.end(), r.reactants.begin(), r.reactants.end());
        entries_.insert(entries_.end(), r.products.begin(), r.products.end());
        productBegin_.push_back(reactantBegin_.back() + (int)r.reactants.size());
        reactantBegin_.push_back(productBegin_.back() + (int)r.products.size());
    }
    Ref operator[](int i) const
    {
        const auto* p = entries_.data();
        return {{p + reactantBegin_[i], p + productBegin_[i]}, {p + productBegin_[i], p + reactantBegin_[i + 1]}};
    }
    size_t size() const { return productBegin_.size(); }

private:
    std::vector<int>                       reactantBegin_{0}, productBegin_{};
    std::vector<Reaction::ReactantProduct> entries_;
};

/* ---- functor tags (the parameterization runs on the device) ---- */
template<int KIND>
struct ComputeReactionRatesFunctor
{
    static constexpr int kind = KIND;
};
namespace net14
{
using ComputeReactionRatesFunctor = nnet_amd::ComputeReactionRatesFunctor<NNET_RATES_NET14>;
}
namespace net86
{
using ComputeReactionRatesFunctor = nnet_amd::ComputeReactionRatesFunctor<NNET_RATES_NET86>;
}
namespace net87
{
using ComputeReactionRatesFunctor = nnet_amd::ComputeReactionRatesFunctor<NNET_RATES_NET87>;
}

namespace eos
{
/* ideal_gas.hpp:45-80 */
struct IdealGasFunctor
{
    double mu;
    explicit IdealGasFunctor(double mu_)
        : mu(mu_)
    {
    }
};
/* helmholtz.hpp:149-200: the species charges (dimension = their number) */
struct HelmholtzFunctor
{
    std::vector<double> Z;
    template<class Vector>
    explicit HelmholtzFunctor(const Vector& Z_)
        : Z(std::begin(Z_), std::end(Z_))
    {
    }
    template<class Vector>
    HelmholtzFunctor(const Vector& Z_, int dimension)
        : Z(std::begin(Z_), std::begin(Z_) + dimension)
    {
    }
};
} // namespace eos

/* ---- error type of this interface (the C-ABI returns codes) ---- */
class Error : public std::runtime_error
{
public:
    Error(int code, const std::string& what)
        : std::runtime_error(what)
        , code_(code)
    {
    }
    int code() const { return code_; }

private:
    int code_;
};

/* ---- functor traits ----
 * RatesKind<T>::kind(f): the rate parameterization of functor f (NNET_RATES_NET14/86/87); throws
 * Error(NNET_E_ARG) for a functor the device does not implement.
 * EosDesc<T>::describe(e, b): fills b.eos and b.mu (ideal gas) or b.Z and b.z_dim (Helmholtz; Z
 * is read during the call only). */
template<class T, class Enable = void>
struct RatesKind;
template<int K>
struct RatesKind<ComputeReactionRatesFunctor<K>>
{
    static int kind(const ComputeReactionRatesFunctor<K>&) { return K; }
};

template<class T, class Enable = void>
struct EosDesc;
template<>
struct EosDesc<eos::IdealGasFunctor>
{
    static void describe(const eos::IdealGasFunctor& e, nnet_batch& b)
    {
        b.eos = NNET_EOS_IDEAL;
        b.mu  = e.mu;
    }
};
template<>
struct EosDesc<eos::HelmholtzFunctor>
{
    static void describe(const eos::HelmholtzFunctor& e, nnet_batch& b)
    {
        b.eos   = NNET_EOS_HELMHOLTZ;
        b.Z     = e.Z.data();
        b.z_dim = (int)e.Z.size();
    }
};

namespace detail
{
template<class P>
double* raw(P p)
{
    if constexpr (std::is_pointer_v<P>) return const_cast<double*>(reinterpret_cast<const double*>(p));
    else
        return reinterpret_cast<double*>(p.get());
}

/* Is n.devData.rho_m1 allocated over [0, lastIndex)?  Like the reference
 * (parallel_nuclear_net.hpp:99-112): through fieldNames / isAllocated(index) when the Data type
 * has them, else by the array's size. */
template<class Dev>
auto rho_m1_allocated(const Dev& dev, size_t lastIndex, int) -> decltype(dev.isAllocated(0), dev.fieldNames.begin(), bool())
{
    size_t idx = 0;
    for (auto it = dev.fieldNames.begin(); it != dev.fieldNames.end() && std::string(*it) != "rho_m1"; ++it)
        ++idx;
    return dev.isAllocated(idx) && dev.rho_m1.size() >= lastIndex;
}
template<class Dev>
bool rho_m1_allocated(const Dev& dev, size_t lastIndex, long)
{
    return dev.rho_m1.size() >= lastIndex && dev.rho_m1.size() > 0;
}

inline void check(nnet_ctx* ctx, int rc, const char* where)
{
    if (rc != NNET_OK) throw Error(rc, std::string(where) + ": " + (ctx ? nnet_last_error(ctx) : "error"));
}

/* One context per device for the process, created on first use.  The Helmholtz table path is
 * taken from set_helmholtz_table() or, failing that, the NNET_HELM_TABLE environment variable. */
struct Registry
{
    std::mutex                                     mu;
    std::map<int, nnet_ctx*>                       ctx;
    std::string                                    helm_path;
    std::map<std::pair<int, std::vector<int>>, nnet_network*> nets; // (device, kind + list) -> network

    static Registry& get()
    {
        static Registry r;
        return r;
    }
    nnet_ctx* context(int device)
    {
        auto it = ctx.find(device);
        if (it != ctx.end()) return it->second;
        const char* path = helm_path.empty() ? std::getenv("NNET_HELM_TABLE") : helm_path.c_str();
        nnet_ctx*   c    = nullptr;
        const int   rc   = nnet_ctx_create(&c, device, path);
        if (rc != NNET_OK) throw Error(rc, "nnet_ctx_create failed" + std::string(c ? ": " : "") + (c ? nnet_last_error(c) : ""));
        ctx[device] = c;
        return c;
    }
};

template<class List>
nnet_network* network(nnet_ctx* c, int device, int kind, const List& reactions)
{
    std::vector<int> rb{0}, pb, ids, nums;
    for (size_t r = 0; r < reactions.size(); ++r)
    {
        auto ref = reactions[(int)r];
        for (auto& x : ref.reactants)
            ids.push_back(x.speciesId), nums.push_back(x.numConsumed);
        pb.push_back((int)ids.size());
        for (auto& x : ref.products)
            ids.push_back(x.speciesId), nums.push_back(x.numConsumed);
        rb.push_back((int)ids.size());
    }
    std::vector<int> key{kind};
    key.insert(key.end(), rb.begin(), rb.end());
    key.insert(key.end(), pb.begin(), pb.end());
    key.insert(key.end(), ids.begin(), ids.end());
    key.insert(key.end(), nums.begin(), nums.end());
    auto& reg = Registry::get();
    auto  it  = reg.nets.find({device, key});
    if (it != reg.nets.end()) return it->second;
    int n_species = 0;
    for (int id : ids)
        n_species = id + 1 > n_species ? id + 1 : n_species;
    nnet_network* net = nullptr;
    check(c,
          nnet_network_create(c, kind, n_species, (int)pb.size(), rb.data(), pb.data(), ids.data(), nums.data(), &net),
          "nnet_network_create");
    reg.nets[{device, key}] = net;
    return net;
}
} // namespace detail

/* Helmholtz table used by contexts created after this call (binary [21][541][201] doubles). */
inline void set_helmholtz_table(const std::string& path)
{
    auto&                       reg = detail::Registry::get();
    std::lock_guard<std::mutex> lk(reg.mu);
    reg.helm_path = path;
    for (auto& kv : reg.ctx)
        detail::check(kv.second, nnet_ctx_load_helmholtz(kv.second, path.c_str()), "nnet_ctx_load_helmholtz");
}

/* Newton-iteration cap per particle and call on `device` (default 10000, 0 = none). */
inline void set_max_iterations(long max_iter, int device = 0)
{
    auto&                       reg = detail::Registry::get();
    std::lock_guard<std::mutex> lk(reg.mu);
    nnet_ctx*                   c = reg.context(device);
    detail::check(c, nnet_ctx_set_max_iterations(c, max_iter), "nnet_ctx_set_max_iterations");
}

/* The library's copies of net14::reactionList, net86::reactionList and net87::reactionList. */
inline ReactionList builtin_reaction_list(int kind, int device = 0)
{
    auto&                       reg = detail::Registry::get();
    std::lock_guard<std::mutex> lk(reg.mu);
    nnet_ctx*                   c   = reg.context(device);
    nnet_network*               net = nullptr;
    detail::check(c, nnet_network_builtin(c, kind, &net), "nnet_network_builtin");
    int nr = 0, ne = 0;
    detail::check(c, nnet_network_reactions(net, &nr, &ne, nullptr, nullptr, nullptr, nullptr), "nnet_network_reactions");
    std::vector<int> rb(nr + 1), pb(nr), ids(ne), nums(ne);
    detail::check(c, nnet_network_reactions(net, &nr, &ne, rb.data(), pb.data(), ids.data(), nums.data()),
                  "nnet_network_reactions");
    nnet_network_destroy(net);
    ReactionList list;
    for (int r = 0; r < nr; ++r)
    {
        Reaction x;
        for (int q = rb[r]; q < pb[r]; ++q)
            x.reactants.push_back({ids[q], nums[q]});
        for (int q = pb[r]; q < rb[r + 1]; ++q)
            x.products.push_back({ids[q], nums[q]});
        list.pushBack(x);
    }
    return list;
}

namespace parallel
{
/* Solve options beyond the reference's signature (defaults reproduce it). */
struct Options
{
    int       device    = 0;
    int       precision = NNET_PRECISION_FP64;
    uint8_t*  status    = nullptr; // device, lastIndex - firstIndex entries (optional)
    uint32_t* substeps  = nullptr; // device (optional)
    uint32_t* newton    = nullptr; // device (optional)
    uint32_t* fp64_solves = nullptr; // device (optional, mixed precision)
    const uint32_t* cost_hint = nullptr; // device (optional): expected cost per particle, e.g. newton of the last call
    const uint32_t* fp64_hint = nullptr; // device (optional, mixed precision): nonzero -> FP64 path, e.g. fp64_solves of the last call
    // what use_drhodt = true means: NNET_DRHODT_INTEGRATE (default: the reference's GPU branch,
    // parallel_nuclear_net.cu:79-81, which this use_gpu interface replaces -- drho/dt drives
    // rho(t), rho_m1 unchanged) or NNET_DRHODT_CPU (the reference CPU path,
    // parallel_nuclear_net_impl.cpp:89-92 -- rho_m1 overwritten by drho/dt, integration with
    // drho/dt = 0)
    int drhodt_mode = NNET_DRHODT_INTEGRATE;
};

template<bool use_gpu, class Data, typename Float, class List, class Rates, class Eos, class nseFunction = void*>
void computeNuclearReactions(Data& n, size_t firstIndex, size_t lastIndex, const Float hydro_dt,
                             const Float previous_dt, const List& reactions, const Rates& construct_rates_BE,
                             const Eos& eos, bool use_drhodt, const nseFunction jumpToNse = nullptr,
                             const Options& opt = Options{})
{
    static_assert(use_gpu, "this implementation runs on the GPU; the reference's CPU branch is not provided");
    static_assert(std::is_same_v<Float, double>, "the rate and EOS functors are FP64");
    (void)jumpToNse;
    if (lastIndex <= firstIndex) return;
    const int                   kind = RatesKind<Rates>::kind(construct_rates_BE);
    auto&                       reg  = detail::Registry::get();
    std::lock_guard<std::mutex> lk(reg.mu);
    nnet_ctx*                   c    = reg.context(opt.device);
    nnet_network*               net  = detail::network(c, opt.device, kind, reactions);

    auto&               dev = n.devData;
    const int           dimension = (int)n.numSpecies;
    std::vector<double*> Y(dimension);
    for (int i = 0; i < dimension; ++i)
        Y[i] = detail::raw(dev.Y[i].data()) + firstIndex;

    nnet_batch b{};
    b.network   = net;
    b.precision = opt.precision;
    b.n         = lastIndex - firstIndex;
    b.rho       = detail::raw(dev.rho.data()) + firstIndex;
    // parallel_nuclear_net.hpp:99-116: drho/dt is switched off, with the reference's message,
    // when rho_m1 is not allocated; the pointer is formed only for an allocated array
    if (use_drhodt && !detail::rho_m1_allocated(dev, lastIndex, 0))
    {
        use_drhodt = false;
        std::cerr << "disabeling using drho/dt because 'rho_m1' isn't alocated !\n";
    }
    b.rho_m1    = use_drhodt ? detail::raw(dev.rho_m1.data()) + firstIndex : nullptr;
    b.Y         = Y.data();
    b.temp      = detail::raw(dev.temp.data()) + firstIndex;
    b.dt        = detail::raw(dev.dt.data()) + firstIndex;
    b.status    = opt.status;
    b.substeps  = opt.substeps;
    b.newton    = opt.newton;
    b.fp64_solves = opt.fp64_solves;
    b.cost_hint   = opt.cost_hint;
    b.fp64_hint   = opt.fp64_hint;
    EosDesc<Eos>::describe(eos, b);
    detail::check(c, nnet_compute_sync(c, &b, 1, hydro_dt, previous_dt, use_drhodt ? opt.drhodt_mode : NNET_DRHODT_OFF),
                  "nnet_compute");
}

/* Post-step EOS refresh, mirroring
 *   template<bool use_gpu, class Data, class Vector>
 *   void nnet::parallel::computeHelmEOS(Data& n, size_t firstIndex, size_t lastIndex, const Vector& Z)
 *                                       (include/nnet/parallel/parallel_nuclear_net.hpp:191-261):
 * fields Y, temp, rho (read) and u, cv, p, c, dpdT (written) of n.devData over
 * [firstIndex, lastIndex), with Z.size() >= n.numSpecies charges.  Synchronises before returning. */
template<bool use_gpu, class Data, class Vector>
void computeHelmEOS(Data& n, size_t firstIndex, size_t lastIndex, const Vector& Z, const Options& opt = Options{})
{
    static_assert(use_gpu, "this implementation runs on the GPU; the reference's CPU branch is not provided");
    if (lastIndex <= firstIndex) return;
    auto&                       reg = detail::Registry::get();
    std::lock_guard<std::mutex> lk(reg.mu);
    nnet_ctx*                   c   = reg.context(opt.device);

    auto&                     dev       = n.devData;
    const int                 dimension = (int)n.numSpecies;
    std::vector<const double*> Y(dimension);
    for (int i = 0; i < dimension; ++i)
        Y[i] = detail::raw(dev.Y[i].data()) + firstIndex;
    std::vector<double> Zd(Z.begin(), Z.begin() + dimension);
    auto at = [&](auto& field) { return detail::raw(field.data()) + firstIndex; };
    detail::check(c,
                  nnet_helm_eos_sync(c, lastIndex - firstIndex, dimension, Zd.data(), at(dev.temp), at(dev.rho),
                                     Y.data(), at(dev.u), at(dev.cv), at(dev.p), at(dev.c), at(dev.dpdT)),
                  "nnet_helm_eos");
}
} // namespace parallel
} // namespace nnet_amd
