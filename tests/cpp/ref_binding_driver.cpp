// The reference's own objects driving the MI355X solve through nnet_amd/reference_binding.hpp:
// nnet::ReactionList, nnet::net14/86/87::computeReactionRates and nnet::eos::IdealGasFunctor /
// HelmholtzFunctor, passed (also as their nnet:: base classes) to
// nnet_amd::parallel::computeNuclearReactions<true>, the drop-in for
// nnet::parallel::computeNuclearReactions<true> (parallel_nuclear_net.hpp:86-172).
//
//   ref_binding_driver map
//       no device: prints the parameterization every reference functor maps to,
//       "map <rates> <eos> <kind> <eos kind> <mu> <z_dim>", and "unknown ... rejected" lines.
//   ref_binding_driver run <net 14|86|87> <eos 0|1> <in.bin> <out.bin> <helm_table.bin>
//       the solve on the device, files as for shim_driver (in: int64 n, double hydro_dt,
//       double mu, rho[n], temp[n], dt[n], Y[d][n], Z[d]; out: temp, dt, Y, status, substeps,
//       newton).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "nnet_amd/reference_binding.hpp"

namespace
{
template<class T>
struct DeviceArray
{
    T*     p = nullptr;
    size_t n = 0;
    explicit DeviceArray(size_t n_ = 0)
        : n(n_)
    {
        if (n && hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc");
    }
    DeviceArray(const DeviceArray&)            = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept
        : p(o.p)
        , n(o.n)
    {
        o.p = nullptr;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        return *this;
    }
    ~DeviceArray()
    {
        if (p) (void)hipFree(p);
    }
    T*     data() const { return p; }
    size_t size() const { return n; }
    void   upload(const T* h) { (void)hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice); }
    void   download(T* h) const { (void)hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost); }
};

// SPH-EXA's device-side field set with its allocation query (parallel_nuclear_net.hpp:76-84)
struct DeviceData
{
    std::vector<std::string>         fieldNames{"rho", "rho_m1", "temp", "dt"};
    DeviceArray<double>              rho, rho_m1, temp, dt;
    std::vector<DeviceArray<double>> Y;
    bool isAllocated(size_t idx) const { return idx < fieldNames.size() && (idx != 1 || rho_m1.size() > 0); }
};
struct ParticleData
{
    int        numSpecies = 0;
    DeviceData devData;
};

struct UnknownRates : nnet::ComputeReactionRatesFunctor<double>
{
};
struct UnknownEos : nnet::EosFunctor<double>
{
};

const nnet::ReactionList& reactions_of(int net)
{
    return net == 14 ? nnet::net14::reactionList : net == 86 ? nnet::net86::reactionList : nnet::net87::reactionList;
}
const nnet::ComputeReactionRatesFunctor<double>& rates_of(int net)
{
    if (net == 14) return nnet::net14::computeReactionRates;
    if (net == 86) return nnet::net86::computeReactionRates;
    return nnet::net87::computeReactionRates;
}

int map_mode()
{
    const std::vector<double> Z(87, 6.0);
    const nnet::eos::IdealGasFunctor<double>  ideal(10.0);
    for (int net : {14, 86, 87})
    {
        const nnet::eos::HelmholtzFunctor<double> helm(Z, net);
        const nnet::EosFunctor<double>*           eos[2] = {&ideal, &helm};
        for (int e = 0; e < 2; ++e)
        {
            nnet_batch b{};
            // through the base classes, as the reference's signature passes them
            const int kind = nnet_amd::RatesKind<nnet::ComputeReactionRatesFunctor<double>>::kind(rates_of(net));
            nnet_amd::EosDesc<nnet::EosFunctor<double>>::describe(*eos[e], b);
            std::printf("map net%d %s %d %d %g %d\n", net, e ? "helmholtz" : "ideal", kind, b.eos, b.eos ? 0.0 : b.mu,
                        b.z_dim);
        }
    }
    // and through the derived types, as a caller holding the concrete objects passes them
    std::printf("derived %d %d %d\n", nnet_amd::RatesKind<nnet::net14::ComputeReactionRatesFunctor<double>>::kind(
                                           nnet::net14::computeReactionRates),
                nnet_amd::RatesKind<nnet::net86::ComputeReactionRatesFunctor<double>>::kind(
                    nnet::net86::computeReactionRates),
                nnet_amd::RatesKind<nnet::net87::ComputeReactionRatesFunctor<double>>::kind(
                    nnet::net87::computeReactionRates));
    try
    {
        nnet_amd::RatesKind<UnknownRates>::kind(UnknownRates{});
        std::printf("unknown rates accepted\n");
    }
    catch (const nnet_amd::Error& err)
    {
        std::printf("unknown rates rejected %d\n", err.code());
    }
    try
    {
        nnet_batch b{};
        nnet_amd::EosDesc<UnknownEos>::describe(UnknownEos{}, b);
        std::printf("unknown eos accepted\n");
    }
    catch (const nnet_amd::Error& err)
    {
        std::printf("unknown eos rejected %d\n", err.code());
    }
    return 0;
}

int run_mode(int argc, char** argv)
{
    if (argc < 7) return 2;
    const int net = std::atoi(argv[2]), eos = std::atoi(argv[3]);
    nnet_amd::set_helmholtz_table(argv[6]);
    FILE* f = std::fopen(argv[4], "rb");
    if (!f) return 3;
    int64_t n;
    double  hdt, mu;
    if (std::fread(&n, 8, 1, f) != 1 || std::fread(&hdt, 8, 1, f) != 1 || std::fread(&mu, 8, 1, f) != 1) return 4;
    const int           d = net;
    std::vector<double> rho(n), temp(n), dt(n), Y((size_t)d * n), Z(d);
    size_t ok = std::fread(rho.data(), 8, n, f) + std::fread(temp.data(), 8, n, f) + std::fread(dt.data(), 8, n, f) +
                std::fread(Y.data(), 8, (size_t)d * n, f) + std::fread(Z.data(), 8, d, f);
    std::fclose(f);
    if (ok != (size_t)(3 * n + (size_t)d * n + d)) return 5;

    ParticleData pd;
    pd.numSpecies = d;
    auto& dv      = pd.devData;
    dv.rho        = DeviceArray<double>(n);
    dv.rho_m1     = DeviceArray<double>(n);
    dv.temp       = DeviceArray<double>(n);
    dv.dt         = DeviceArray<double>(n);
    dv.rho.upload(rho.data());
    dv.rho_m1.upload(rho.data());
    dv.temp.upload(temp.data());
    dv.dt.upload(dt.data());
    for (int s = 0; s < d; ++s)
    {
        dv.Y.emplace_back(n);
        dv.Y.back().upload(Y.data() + (size_t)s * n);
    }
    DeviceArray<uint8_t>        st(n);
    DeviceArray<uint32_t>       sub(n), newton(n);
    nnet_amd::parallel::Options opt;
    opt.status   = st.data();
    opt.substeps = sub.data();
    opt.newton   = newton.data();
    try
    {
        const nnet::eos::IdealGasFunctor<double>  ideal(mu);
        const nnet::eos::HelmholtzFunctor<double> helm(Z, d);
        const nnet::EosFunctor<double>&           e = eos == 1 ? static_cast<const nnet::EosFunctor<double>&>(helm)
                                                               : static_cast<const nnet::EosFunctor<double>&>(ideal);
        nnet_amd::parallel::computeNuclearReactions<true>(pd, 0, (size_t)n, hdt, hdt, reactions_of(net),
                                                          rates_of(net), e, false, nullptr, opt);
    }
    catch (const nnet_amd::Error& err)
    {
        std::fprintf(stderr, "nnet_amd error %d: %s\n", err.code(), err.what());
        return 6;
    }
    dv.temp.download(temp.data());
    dv.dt.download(dt.data());
    for (int s = 0; s < d; ++s)
        dv.Y[s].download(Y.data() + (size_t)s * n);
    std::vector<uint8_t>  hs(n);
    std::vector<uint32_t> hsub(n), hnew(n);
    st.download(hs.data());
    sub.download(hsub.data());
    newton.download(hnew.data());
    FILE* o = std::fopen(argv[5], "wb");
    if (!o) return 7;
    std::fwrite(temp.data(), 8, n, o);
    std::fwrite(dt.data(), 8, n, o);
    std::fwrite(Y.data(), 8, (size_t)d * n, o);
    std::fwrite(hs.data(), 1, n, o);
    std::fwrite(hsub.data(), 4, n, o);
    std::fwrite(hnew.data(), 4, n, o);
    std::fclose(o);
    std::printf("ref_binding_driver: net%d eos %d n=%lld done\n", net, eos, (long long)n);
    return 0;
}
} // namespace

int main(int argc, char** argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "map") == 0) return map_mode();
    if (argc >= 2 && std::strcmp(argv[1], "run") == 0) return run_mode(argc, argv);
    return 2;
}
