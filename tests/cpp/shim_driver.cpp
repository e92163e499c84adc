// Drives nnet_amd::parallel::computeNuclearReactions and, for the Helmholtz EOS, the post-step
// nnet_amd::parallel::computeHelmEOS (the C++ mirrors of the reference's nnet::parallel
// functions of the same names) the way SPH-EXA calls them: a Data object whose devData holds
// device arrays rho, rho_m1, temp, dt, u, cv, p, c, dpdT and Y[numSpecies].
//
//   shim_driver <net 14|86|87> <eos 0|1> <in.bin> <out.bin> [helm_table.bin]
//
// in.bin : int64 n, double hydro_dt, double mu, then rho[n], temp[n], dt[n], Y[d][n],
//          and for the Helmholtz EOS Z[d]
// out.bin: temp[n], dt[n], Y[d][n], status[n] (uint8), substeps[n] (uint32), newton[n] (uint32),
//          and for the Helmholtz EOS u[n], cv[n], p[n], c[n], dpdT[n] after the solve
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "nnet_amd/parallel_nuclear_net.hpp"

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
    T*     data() { return p; }
    size_t size() const { return n; }
    void   upload(const T* h) { (void)hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice); }
    void   download(T* h) const { (void)hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost); }
};

struct DeviceData
{
    DeviceArray<double>              rho, rho_m1, temp, dt, u, cv, p, c, dpdT;
    std::vector<DeviceArray<double>> Y;
};
struct ParticleData
{
    int        numSpecies = 0;
    DeviceData devData;
};

template<class Rates, class Eos>
static void run(ParticleData& n, size_t count, double hydro_dt, const Rates& rates, const Eos& eos, int kind,
                const nnet_amd::parallel::Options& opt)
{
    const auto reactions = nnet_amd::builtin_reaction_list(kind);
    nnet_amd::parallel::computeNuclearReactions<true>(n, 0, count, hydro_dt, hydro_dt, reactions, rates, eos, false,
                                                      nullptr, opt);
}

int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    const int net = std::atoi(argv[1]), eos = std::atoi(argv[2]);
    if (argc > 5) nnet_amd::set_helmholtz_table(argv[5]);
    FILE* f = std::fopen(argv[3], "rb");
    if (!f) return 3;
    int64_t n;
    double  hdt, mu;
    if (std::fread(&n, 8, 1, f) != 1 || std::fread(&hdt, 8, 1, f) != 1 || std::fread(&mu, 8, 1, f) != 1) return 4;
    const int           d = net == 87 ? 87 : net;
    std::vector<double> rho(n), temp(n), dt(n), Y((size_t)d * n), Z(d);
    size_t              ok = std::fread(rho.data(), 8, n, f) + std::fread(temp.data(), 8, n, f) +
                std::fread(dt.data(), 8, n, f) + std::fread(Y.data(), 8, (size_t)d * n, f);
    if (eos == 1) ok += std::fread(Z.data(), 8, d, f);
    std::fclose(f);
    if (ok != (size_t)(3 * n + (size_t)d * n + (eos == 1 ? d : 0))) return 5;

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
    for (auto* a : {&dv.u, &dv.cv, &dv.p, &dv.c, &dv.dpdT})
        *a = DeviceArray<double>(n);
    DeviceArray<uint8_t>          st(n);
    DeviceArray<uint32_t>         sub(n), newton(n);
    nnet_amd::parallel::Options   opt;
    opt.status   = st.data();
    opt.substeps = sub.data();
    opt.newton   = newton.data();
    try
    {
        if (eos == 1)
        {
            const nnet_amd::eos::HelmholtzFunctor helm(Z);
            if (net == 14) run(pd, n, hdt, nnet_amd::net14::ComputeReactionRatesFunctor{}, helm, 14, opt);
            else if (net == 86) run(pd, n, hdt, nnet_amd::net86::ComputeReactionRatesFunctor{}, helm, 86, opt);
            else run(pd, n, hdt, nnet_amd::net87::ComputeReactionRatesFunctor{}, helm, 87, opt);
            nnet_amd::parallel::computeHelmEOS<true>(pd, 0, n, Z);
        }
        else
        {
            const nnet_amd::eos::IdealGasFunctor ideal(mu);
            if (net == 14) run(pd, n, hdt, nnet_amd::net14::ComputeReactionRatesFunctor{}, ideal, 14, opt);
            else if (net == 86) run(pd, n, hdt, nnet_amd::net86::ComputeReactionRatesFunctor{}, ideal, 86, opt);
            else run(pd, n, hdt, nnet_amd::net87::ComputeReactionRatesFunctor{}, ideal, 87, opt);
        }
    }
    catch (const nnet_amd::Error& e)
    {
        std::fprintf(stderr, "nnet_amd error %d: %s\n", e.code(), e.what());
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
    FILE* o = std::fopen(argv[4], "wb");
    if (!o) return 7;
    std::fwrite(temp.data(), 8, n, o);
    std::fwrite(dt.data(), 8, n, o);
    std::fwrite(Y.data(), 8, (size_t)d * n, o);
    std::fwrite(hs.data(), 1, n, o);
    std::fwrite(hsub.data(), 4, n, o);
    std::fwrite(hnew.data(), 4, n, o);
    if (eos == 1)
        for (auto* a : {&dv.u, &dv.cv, &dv.p, &dv.c, &dv.dpdT})
        {
            std::vector<double> h(n);
            a->download(h.data());
            std::fwrite(h.data(), 8, n, o);
        }
    std::fclose(o);
    std::printf("shim_driver: net%d eos %d n=%lld done\n", net, eos, (long long)n);
    return 0;
}
