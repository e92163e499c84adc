// TEST INFRASTRUCTURE — the CPU oracle.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load liboracle.so, and only as the checker (or, for cpu_baseline, as the
// timed CPU reference).  The product library never links or calls it.
//
// A plain-C++ FP64 restatement of SPH-EXA-nnet's CPU batch path
//   nnet::parallel::computeNuclearReactions<false>   parallel/parallel_nuclear_net.hpp:86-172
//   computeNuclearReactionsImpl                      parallel/implementation/parallel_nuclear_net_impl.cpp:63-103
//   solveSystemSubstep and the Newton/substep loop   nuclear_net.hpp:692-959
//   prepareSystemFromGuess (assembly)                nuclear_net.hpp:514-609
//   derivativesFromReactions / firstOrderDYFromReactions  nuclear_net.hpp:347-438
//   eigen::solve                                     nnet_util/eigen.hpp:179-223
//   reaction lists                                   net14.hpp:69-134, net86.hpp:95-192, net87.hpp:88-98
// with the same operation order, plus three things the reference lacks: per-particle counters
// (accepted substeps, Newton iterations, rejections), exact detection of the controller's
// infinite replay (a rejected attempt whose dt and Newton seed come back bit-identical is
// replayed forever by the reference; the particle is stopped and flagged STALLED_REPLAY), and a
// per-particle status instead of the host exception on dt == 0.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp, no -march: the same code
// generation as the reference build in oracle/_ref, so results are bit-identical).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <omp.h>

#include "oracle_helmholtz.hpp"
#include "oracle_rates.hpp"

namespace oracle
{
// nuclear_net.hpp:48-106
namespace k
{
const double theta = 0.8, minTemp = 1e8, minRho = 1e5, maxDtStep = 2, minDtStep = 1e-2, nanDtStep = 2e-1;
const double epsilonSystem = 1e-40, epsilonVector = 1e-16;
const double NR_maxDt = 1e-2, NR_dTOverTempTarget = 1e-2, NR_dTOverTempTol = 4, NR_itTol = 1e-7;
const int    NR_minIt = 1, NR_maxIt = 11;
const double substep_dtTol = 1e-6;
} // namespace k

enum Status : uint8_t
{
    ST_OK = 0, ST_SKIPPED = 1, ST_STALLED = 2, ST_ZERO_DT = 3, ST_MAX_ITER = 4
};

// ReactionList in the reference's flat layout (nuclear_net.hpp:168-213).
struct Reactions
{
    std::vector<int> reactantBegin{0}, productBegin, id, num;
    void push(std::initializer_list<std::pair<int, int>> r, std::initializer_list<std::pair<int, int>> p)
    {
        for (auto& x : r)
            id.push_back(x.first), num.push_back(x.second);
        for (auto& x : p)
            id.push_back(x.first), num.push_back(x.second);
        productBegin.push_back(reactantBegin.back() + (int)r.size());
        reactantBegin.push_back(productBegin.back() + (int)p.size());
    }
    int size() const { return (int)productBegin.size(); }
};

static Reactions make_net14()
{
    Reactions r;
    r.push({{0, 3}}, {{1, 1}});
    r.push({{0, 1}, {1, 1}}, {{2, 1}});
    r.push({{0, 1}, {2, 1}}, {{3, 1}});
    for (int s = 3; s < 13; ++s)
        r.push({{0, 1}, {s, 1}}, {{s + 1, 1}});
    r.push({{1, 2}}, {{3, 1}, {0, 1}});
    r.push({{1, 1}, {2, 1}}, {{4, 1}, {0, 1}});
    r.push({{2, 2}}, {{5, 1}, {0, 1}});
    r.push({{1, 1}}, {{0, 3}});
    r.push({{2, 1}}, {{0, 1}, {1, 1}});
    r.push({{3, 1}}, {{0, 1}, {2, 1}});
    for (int s = 4; s < 14; ++s)
        r.push({{s, 1}}, {{0, 1}, {s - 1, 1}});
    return r;
}

static Reactions make_net86()
{
    const int* mr = NNET86_MAIN_REACTANT;
    const int* sr = NNET86_SECONDARY_REACTANT;
    const int* mp = NNET86_MAIN_PRODUCT;
    const int  proton = 0, neutron = 1, alpha = 2;
    Reactions  r;
    r.push({{mr[0], 2}}, {{mp[0], 1}, {alpha, 1}});
    r.push({{mr[1], 1}, {sr[1], 1}}, {{mp[1], 1}, {alpha, 1}});
    r.push({{mr[2], 2}}, {{mp[2], 1}, {alpha, 1}});
    r.push({{mr[4], 3}}, {{mp[4], 1}});
    for (int i = 5; i < 157; ++i)
    {
        int dZ = (int)(NNET86_Z[mr[i]] + NNET86_Z[sr[i]] - NNET86_Z[mp[i]]);
        int dA = (int)(NNET86_A[mr[i]] + NNET86_A[sr[i]] - NNET86_A[mp[i]]);
        if (dZ == 0 && dA == 0) r.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}});
        else if (dA == 1 && dZ == 0) r.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}, {neutron, 1}});
        else if (dA == 1 && dZ == 1) r.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}, {proton, 1}});
        else if (dA == 4 && dZ == 2) r.push({{mr[i], 1}, {sr[i], 1}}, {{mp[i], 1}, {alpha, 1}});
        else
            return Reactions{};
    }
    r.push({{mp[4], 1}}, {{mr[4], 3}});
    for (int i = 5; i < 157; ++i)
    {
        int dZ = (int)(NNET86_Z[mp[i]] - (NNET86_Z[mr[i]] + NNET86_Z[sr[i]]));
        int dA = (int)(NNET86_A[mp[i]] - (NNET86_A[mr[i]] + NNET86_A[sr[i]]));
        if (dZ == 0 && dA == 0) r.push({{mp[i], 1}}, {{mr[i], 1}, {sr[i], 1}});
        else if (dA == -1 && dZ == 0) r.push({{mp[i], 1}, {neutron, 1}}, {{mr[i], 1}, {sr[i], 1}});
        else if (dA == -1 && dZ == -1) r.push({{mp[i], 1}, {proton, 1}}, {{mr[i], 1}, {sr[i], 1}});
        else if (dA == -4 && dZ == -2) r.push({{mp[i], 1}, {alpha, 1}}, {{mr[i], 1}, {sr[i], 1}});
        else
            return Reactions{};
    }
    return r;
}

static Reactions make_net87()
{
    Reactions r = make_net86();
    r.push({{0, 1}, {86, 1}}, {{1, 1}});
    r.push({{1, 1}, {86, 1}}, {{0, 1}});
    return r;
}

static const Reactions& reactions_for(int net)
{
    static const Reactions r14 = make_net14(), r86 = make_net86(), r87 = make_net87();
    return net == 14 ? r14 : net == 86 ? r86 : r87;
}

static helm::Table g_helm;

// The EOS functor: ideal gas (eos/ideal_gas.hpp:69-82) or Helmholtz (eos/helmholtz.hpp:217-225).
struct EosSpec
{
    int           kind; // 0 ideal gas, 1 Helmholtz
    double        mu;
    const double* Z;
    int           zdim;
};

static EosOut eval_eos(const EosSpec& e, const double* Y, double T, double rho)
{
    EosOut r{};
    if (e.kind == 0)
    {
        const double dmy = 8.317e7 / e.mu;
        r.cv             = 1.5 * dmy;
        r.u              = T * r.cv;
        r.p              = rho * T * dmy;
        r.c              = std::sqrt(r.p / rho);
        r.dpdT           = rho * dmy;
        return r;
    }
    double abar = 0.;
    for (int s = 0; s < e.zdim; ++s)
        abar += Y[s];
    double zbar = 0;
    for (int s = 0; s < e.zdim; ++s)
        zbar += Y[s] * e.Z[s];
    return helm::eos(g_helm, abar, zbar, T, rho);
}

static void eval_rates(int net, const double* Y, double T, double rho, double* BE, double* rates, double* drates)
{
    if (net == 14) rates_net14(T, rho, BE, rates, drates);
    else if (net == 86) rates_net86(T, rho, BE, rates, drates);
    else
        rates_net87(Y, T, rho, BE, rates, drates);
}

// nuclear_net.hpp:347-381
static void derivatives_from_reactions(const Reactions& R, const double* rates, double rho, const double* Y, double* dY,
                                       int d)
{
    for (int i = 0; i < d; ++i)
        dY[i] = 0.;
    for (int r = 0; r < R.size(); ++r)
    {
        double rate = rates[r];
        for (int q = R.reactantBegin[r]; q < R.productBegin[r]; ++q)
        {
            if (R.num[q] != 1) rate /= std::tgamma(R.num[q] + 1);
            rate *= std::pow(Y[R.id[q]] * rho, R.num[q]);
        }
        rate /= rho;
        for (int q = R.reactantBegin[r]; q < R.productBegin[r]; ++q)
            dY[R.id[q]] -= rate * R.num[q];
        for (int q = R.productBegin[r]; q < R.reactantBegin[r + 1]; ++q)
            dY[R.id[q]] += rate * R.num[q];
    }
}

// nuclear_net.hpp:395-438 (writes the species block of the (d+1)x(d+1) column-major matrix)
static void first_order_dy(const Reactions& R, const double* rates, double rho, const double* Y, double* M, int d)
{
    const int n = d + 1;
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j)
            M[(i + 1) + n * (j + 1)] = 0.;
    for (int r = 0; r < R.size(); ++r)
    {
        double rate = rates[r];
        for (int q = R.reactantBegin[r]; q < R.productBegin[r]; ++q)
        {
            if (R.num[q] != 1) rate /= std::tgamma(R.num[q] + 1);
            rate *= std::pow(Y[R.id[q]] * rho, R.num[q] - 1);
        }
        if (rate > k::epsilonSystem)
            for (int q = R.reactantBegin[r]; q < R.productBegin[r]; ++q)
            {
                const int b    = R.id[q];
                double    this_rate = rate;
                for (int o = R.reactantBegin[r]; o < R.productBegin[r]; ++o)
                    if (R.id[o] != b) this_rate *= Y[R.id[o]] * rho;
                for (int o = R.reactantBegin[r]; o < R.productBegin[r]; ++o)
                    M[(R.id[o] + 1) + n * (b + 1)] -= this_rate * R.num[o];
                for (int o = R.productBegin[r]; o < R.reactantBegin[r + 1]; ++o)
                    M[(R.id[o] + 1) + n * (b + 1)] += this_rate * R.num[o];
            }
    }
}

// nuclear_net.hpp:514-609.  Returns false on dt == 0 (the reference throws there).
static bool prepare_from_guess(int net, const Reactions& R, int d, double* Mp, double* RHS, double* rates,
                               const double* Y, double T, const double* Yg, double Tg, double rho, double drho_dt,
                               const EosOut& eos, double dt)
{
    if (dt == 0) return false;
    const int n         = d + 1;
    double*   BE        = RHS + 1;
    double*   drates_dT = Mp + n;

    eval_rates(net, Yg, Tg, rho, BE, rates, drates_dT);
    derivatives_from_reactions(R, drates_dT, rho, Yg, &Mp[1], d);
    for (int i = 0; i < d; ++i)
        Mp[0 + n * (i + 1)] = -BE[i] / eos.cv;

    derivatives_from_reactions(R, rates, rho, Yg, &RHS[1], d);
    for (int i = 0; i < d; ++i)
        RHS[i + 1] *= dt;
    for (int i = 0; i < d; ++i)
    {
        RHS[i + 1] += -k::theta * dt * Mp[(i + 1) + 0] * (Tg - T);
        Mp[(i + 1) + 0] *= -k::theta * dt;
    }

    first_order_dy(R, rates, rho, Yg, Mp, d);
    for (int i = 0; i < d; ++i)
    {
        double corr = 0;
        for (int j = 0; j < d; ++j)
            corr += Mp[(i + 1) + n * (j + 1)] * (Yg[j] - Y[j]);
        RHS[i + 1] += -k::theta * dt * corr;
    }
    for (int i = 0; i < d; ++i)
    {
        Mp[(i + 1) + n * (i + 1)] = 1. - k::theta * dt * Mp[(i + 1) + n * (i + 1)];
        for (int j = 0; j < d; ++j)
            if (i != j) Mp[(i + 1) + n * (j + 1)] = -k::theta * dt * Mp[(i + 1) + n * (j + 1)];
    }
    const double drho    = drho_dt * dt;
    const double value_1 = eos.dpdT * drho / (rho * rho);
    RHS[0]               = T * value_1 / eos.cv;
    Mp[0]                = 1 - k::theta * value_1 / eos.cv;
    return true;
}

// eigen.hpp:179-223: unpivoted Gaussian elimination to unit upper form + back substitution.
static void gauss_solve(double* M, double* RHS, double* X, int n, double eps)
{
    for (int i = 0; i < n; ++i)
    {
        for (int j = 0; j < i; ++j)
        {
            const double w = M[i + n * j];
            M[i + n * j]   = 0;
            if (std::abs(w) > eps)
            {
                RHS[i] -= w * RHS[j];
                for (int c = j + 1; c < n; ++c)
                    M[i + n * c] -= w * M[j + n * c];
            }
        }
        const double diag = M[i + n * i];
        M[i + n * i]      = 1;
        RHS[i] /= diag;
        for (int j = i + 1; j < n; ++j)
            M[i + n * j] /= diag;
    }
    for (int i = n - 1; i >= 0; --i)
    {
        double res = RHS[i];
        for (int j = i + 1; j < n; ++j)
            res -= M[i + n * j] * X[j];
        X[i] = res;
    }
}

// ---------------------------------------------------------------------------------------------
// Models of mixed-precision linear solves (test infrastructure): the device's mixed path (FP32
// unpivoted LU in the reference's elimination order, FP64 residuals, LU iterative refinement) and
// an alternative weighed against it (GMRES-IR: FP64 GMRES on the Ruiz-equilibrated system, left
// preconditioned by its FP32 LU factors; Carson & Higham, SIAM J. Sci. Comput. 40(2), 2018).
// mode 0 is the reference's FP64 eigen::solve; with mode 1 or 2 an iteration whose correction does
// not converge (every |d_i| <= tol (|x_i| + s_i), s = (T, Y)) abandons the particle, which the batch
// then integrates again in FP64 from the start of the step -- the device's redo pass.
struct SolverModel
{
    int    mode   = 0;
    double tol    = 1e-11;
    int    sweeps = 1; // refinement sweeps (mode 1), outer iterations (mode 2)
    int    inner  = 8; // GMRES iterations per outer iteration (mode 2)
};
static SolverModel& solver_model()
{
    static SolverModel m;
    return m;
}
struct SolverCounts
{
    long solves = 0, converged = 0, inner = 0;
};
static SolverCounts& solver_counts()
{
    static SolverCounts c;
    return c;
}
// set while a batch integrates an abandoned particle again in FP64
static thread_local bool force_fp64 = false;

// unpivoted LU of the column-major n x n matrix A (float), reference order and 1e-40 skip:
// unit-diagonal U above the diagonal, multipliers L below, reciprocal pivots in rinv
static void lu32(std::vector<float>& A, std::vector<float>& rinv, int n)
{
    const float eps = (float)k::epsilonSystem;
    for (int i = 0; i < n; ++i)
    {
        for (int j = 0; j < i; ++j)
        {
            const float w = A[i + n * j];
            if (!(std::abs(w) > eps))
            {
                A[i + n * j] = 0;
                continue;
            }
            for (int c = j + 1; c < n; ++c)
                A[i + n * c] -= w * A[j + n * c];
        }
        rinv[i] = 1.f / A[i + n * i];
        for (int c = i + 1; c < n; ++c)
            A[i + n * c] *= rinv[i];
    }
}
// x = (LU)^-1 b in FP32
static void lu32_solve(const std::vector<float>& A, const std::vector<float>& rinv, int n, const double* b, double* x)
{
    std::vector<float> y(n);
    for (int i = 0; i < n; ++i)
    {
        float v = (float)b[i];
        for (int j = 0; j < i; ++j)
            v -= A[i + n * j] * y[j];
        y[i] = v * rinv[i];
    }
    for (int i = n - 1; i >= 0; --i)
    {
        float v = y[i];
        for (int j = i + 1; j < n; ++j)
            v -= A[i + n * j] * y[j];
        y[i] = v;
    }
    for (int i = 0; i < n; ++i)
        x[i] = (double)y[i];
}
static void residual(const double* M, const double* b, const double* x, double* r, int n)
{
    for (int i = 0; i < n; ++i)
    {
        double v = b[i];
        for (int j = 0; j < n; ++j)
            v = std::fma(-M[i + n * j], x[j], v);
        r[i] = v;
    }
}
static bool small_correction(const double* d, const double* x, const double* s, int n, double tol)
{
    for (int i = 0; i < n; ++i)
        if (!(std::abs(d[i]) <= tol * (std::abs(x[i]) + s[i] + 1e-30))) return false;
    return true;
}

// The mixed solve of M x = b (M column-major, unmodified); s: the magnitude scale of each
// unknown.  Returns false when the correction does not converge.
static bool mixed_solve(const double* M, const double* b, double* x, const double* s, int n)
{
    const SolverModel& m = solver_model();
    std::vector<double> r(n), dx(n), xn(n);
    std::vector<float>  A(n * n), rinv(n);
    if (m.mode == 1)
    {
        for (int i = 0; i < n * n; ++i)
            A[i] = (float)M[i];
        lu32(A, rinv, n);
        lu32_solve(A, rinv, n, b, x);
        for (int sw = 0; sw < m.sweeps; ++sw)
        {
            residual(M, b, x, r.data(), n);
            lu32_solve(A, rinv, n, r.data(), dx.data());
            for (int i = 0; i < n; ++i)
                xn[i] = x[i] + dx[i];
            const bool ok = small_correction(dx.data(), xn.data(), s, n, m.tol);
            std::copy(xn.begin(), xn.end(), x);
            if (ok) return true;
        }
        return false;
    }
    // GMRES-IR on the Ruiz-equilibrated system D_r M D_c (power-of-two scalings)
    std::vector<double> dr(n, 1.), dc(n, 1.), S(M, M + n * n);
    for (int it = 0; it < 8; ++it)
    {
        for (int i = 0; i < n; ++i)
        {
            double a = 0;
            for (int j = 0; j < n; ++j)
                a = std::max(a, std::abs(S[i + n * j]));
            const double f = a > 0 ? std::ldexp(1., -std::ilogb(a) / 2) : 1.;
            dr[i] *= f;
            for (int j = 0; j < n; ++j)
                S[i + n * j] *= f;
        }
        for (int j = 0; j < n; ++j)
        {
            double a = 0;
            for (int i = 0; i < n; ++i)
                a = std::max(a, std::abs(S[i + n * j]));
            const double f = a > 0 ? std::ldexp(1., -std::ilogb(a) / 2) : 1.;
            dc[j] *= f;
            for (int i = 0; i < n; ++i)
                S[i + n * j] *= f;
        }
    }
    for (int i = 0; i < n * n; ++i)
        A[i] = (float)S[i];
    lu32(A, rinv, n);
    auto precond = [&](const double* v, double* out) { lu32_solve(A, rinv, n, v, out); };
    // x0 = D_c (LU)^-1 D_r b
    {
        std::vector<double> t(n);
        for (int i = 0; i < n; ++i)
            t[i] = dr[i] * b[i];
        precond(t.data(), x);
        for (int i = 0; i < n; ++i)
            x[i] *= dc[i];
    }
    const int           K = m.inner;
    std::vector<double> V((size_t)(K + 1) * n), H((size_t)(K + 1) * K, 0.), g(K + 1), cs(K), sn(K), y(K), w(n), t(n);
    for (int outer = 0; outer < m.sweeps; ++outer)
    {
        residual(M, b, x, r.data(), n);
        for (int i = 0; i < n; ++i)
            t[i] = dr[i] * r[i];
        precond(t.data(), w.data()); // z0 = (LU)^-1 D_r r
        double beta = 0;
        for (int i = 0; i < n; ++i)
            beta += w[i] * w[i];
        beta = std::sqrt(beta);
        std::fill(dx.begin(), dx.end(), 0.);
        if (beta > 0)
        {
            for (int i = 0; i < n; ++i)
                V[i] = w[i] / beta;
            std::fill(g.begin(), g.end(), 0.);
            g[0]  = beta;
            int k = 0;
            for (; k < K; ++k)
            {
                // w = (LU)^-1 S v_k
                for (int i = 0; i < n; ++i)
                {
                    double v = 0;
                    for (int j = 0; j < n; ++j)
                        v += S[i + n * j] * V[(size_t)k * n + j];
                    t[i] = v;
                }
                precond(t.data(), w.data());
                for (int q = 0; q <= k; ++q)
                {
                    double h = 0;
                    for (int i = 0; i < n; ++i)
                        h += w[i] * V[(size_t)q * n + i];
                    H[(size_t)q * K + k] = h;
                    for (int i = 0; i < n; ++i)
                        w[i] -= h * V[(size_t)q * n + i];
                }
                double hn = 0;
                for (int i = 0; i < n; ++i)
                    hn += w[i] * w[i];
                hn                         = std::sqrt(hn);
                H[(size_t)(k + 1) * K + k] = hn;
                for (int i = 0; i < n; ++i)
                    V[(size_t)(k + 1) * n + i] = hn > 0 ? w[i] / hn : 0.;
                for (int q = 0; q < k; ++q)
                {
                    const double a = H[(size_t)q * K + k], c = H[(size_t)(q + 1) * K + k];
                    H[(size_t)q * K + k]       = cs[q] * a + sn[q] * c;
                    H[(size_t)(q + 1) * K + k] = -sn[q] * a + cs[q] * c;
                }
                const double a = H[(size_t)k * K + k], c = H[(size_t)(k + 1) * K + k], rr = std::hypot(a, c);
                cs[k]                      = rr > 0 ? a / rr : 1.;
                sn[k]                      = rr > 0 ? c / rr : 0.;
                H[(size_t)k * K + k]       = rr;
                H[(size_t)(k + 1) * K + k] = 0;
                g[k + 1]                   = -sn[k] * g[k];
                g[k]                       = cs[k] * g[k];
#pragma omp atomic
                ++solver_counts().inner;
                if (std::abs(g[k + 1]) <= 1e-8 * beta) {++k; break;}
            }
            for (int q = k - 1; q >= 0; --q)
            {
                double v = g[q];
                for (int p = q + 1; p < k; ++p)
                    v -= H[(size_t)q * K + p] * y[p];
                y[q] = v / H[(size_t)q * K + q];
            }
            for (int i = 0; i < n; ++i)
            {
                double v = 0;
                for (int q = 0; q < k; ++q)
                    v += V[(size_t)q * n + i] * y[q];
                dx[i] = dc[i] * v;
            }
        }
        for (int i = 0; i < n; ++i)
            xn[i] = x[i] + dx[i];
        const bool ok = small_correction(dx.data(), xn.data(), s, n, m.tol);
        std::copy(xn.begin(), xn.end(), x);
        if (ok) return true;
    }
    return false;
}

struct Work
{
    std::vector<double> Mp, RHS, DY, rates, Ybuf;
    void size(int d, int nR)
    {
        Mp.assign((size_t)(d + 1) * (d + 1), 0.);
        RHS.assign(d + 1, 0.);
        DY.assign(d + 1, 0.);
        rates.assign(nR, 0.);
        Ybuf.assign(d, 0.);
    }
};

struct ParticleStats
{
    uint32_t substeps = 0, newton = 0, rejections = 0, max_run = 0, run = 0;
    uint8_t  status   = ST_OK;
    bool     abandoned = false; // a mixed solve did not converge (solver model 1 or 2)
};

// Optional per-iteration trace, for golden-vector fixtures.
struct Trace
{
    int     cap = 0, len = 0;
    double* rec = nullptr; // 6 doubles per Newton iteration: i, used_dt, elapsed, dt_after, T_next, converged
    void    push(int i, double used, double elapsed, double dt, double Tn, int conv)
    {
        if (len < cap)
        {
            double* p = rec + 6 * (size_t)len;
            p[0] = i, p[1] = used, p[2] = elapsed, p[3] = dt, p[4] = Tn, p[5] = conv;
        }
        ++len;
    }
};

static bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// solveSystemSubstep (nuclear_net.hpp:931-959) with prepareSystemSubstep (:850-880),
// prepareSystemNR (:692-717), finalizeSystemNR (:724-782), finalizeSystemSubstep (:887-922).
static ParticleStats solve_particle(int net, int d, const EosSpec& eosSpec, Work& w, double* Y, double& T, double rho,
                                    double drho_dt, double dt_tot, double& dt, long max_iter, Trace* tr)
{
    ParticleStats st;
    if (rho < k::minRho || T < k::minTemp) return st;
    const Reactions& R       = reactions_for(net);
    double*          Ybuf    = w.Ybuf.data();
    double           elapsed = 0;
    // The reference leaves this buffer uninitialised before the first Newton iteration;
    // NaN makes the first convergence test fail, as any value far from the solution does.
    double T_buf        = std::nan("");
    double att_dt       = 0, att_Tbuf = 0;
    long   total_iter   = 0;

    for (int i = 1;; ++i)
    {
        const int it = i;
        if (i == 1)
        {
            att_dt   = dt;
            att_Tbuf = T_buf;
        }
        // prepareSystemSubstep
        const double rho_t   = rho - drho_dt * (dt_tot - elapsed);
        double       used_dt = dt;
        if (dt_tot - elapsed < dt) used_dt = dt_tot - elapsed;
        // prepareSystemNR
        double T_loc = T_buf;
        if (i <= 1)
        {
            for (int j = 0; j < d; ++j)
                Ybuf[j] = Y[j];
            T_loc = T;
        }
        const double T_theta = (1 - k::theta) * T + k::theta * T_loc;
        for (int j = 0; j < d; ++j)
            Ybuf[j] = (1 - k::theta) * Y[j] + k::theta * Ybuf[j];
        const EosOut eos = eval_eos(eosSpec, Ybuf, T_theta, rho_t);
        if (!prepare_from_guess(net, R, d, w.Mp.data(), w.RHS.data(), w.rates.data(), Y, T, Ybuf, T_theta, rho_t,
                                drho_dt, eos, used_dt))
        {
            st.status = ST_ZERO_DT;
            return st;
        }
        if (solver_model().mode == 0 || force_fp64)
            gauss_solve(w.Mp.data(), w.RHS.data(), w.DY.data(), d + 1, k::epsilonSystem);
        else
        {
            std::vector<double> sc(d + 1);
            sc[0] = T;
            for (int j = 0; j < d; ++j)
                sc[j + 1] = std::abs(Y[j]);
            const bool ok = mixed_solve(w.Mp.data(), w.RHS.data(), w.DY.data(), sc.data(), d + 1);
#pragma omp atomic
            ++solver_counts().solves;
            if (!ok)
            {
                st.abandoned = true;
                return st;
            }
#pragma omp atomic
            ++solver_counts().converged;
        }
        ++st.newton;
        ++total_iter;

        // finalizeSystemSubstep / finalizeSystemNR
        double u = dt;
        if (dt_tot - elapsed < dt) u = dt_tot - elapsed;
        double       timestep = 0;
        bool         conv     = false;
        const double last_T   = T_buf;
        for (int j = 0; j < d; ++j)
            Ybuf[j] = Y[j] + w.DY[j + 1];
        T_buf        = T + w.DY[0];
        bool has_nan = std::isnan(T_buf);
        for (int j = 0; j < d && !has_nan; ++j)
            has_nan = std::isnan(Ybuf[j]);
        if (has_nan || T_buf < 0)
        {
            u *= k::nanDtStep;
            i        = 0;
            timestep = 0.;
        }
        else
        {
            const double dT_T = std::abs((T_buf - T) / T_buf);
            if (i >= k::NR_minIt && dT_T > k::NR_dTOverTempTarget * k::NR_dTOverTempTol)
            {
                u *= k::NR_dTOverTempTol / dT_T;
                i        = 0;
                timestep = 0.;
                for (int j = 0; j < d; ++j)
                    Ybuf[j] = Y[j];
                T_buf = T;
            }
            else
            {
                for (int j = 0; j < d; ++j)
                    if (Ybuf[j] <= k::epsilonVector) Ybuf[j] = 0;
                const double correction = std::abs((T_buf - last_T) / T_buf);
                if ((i >= k::NR_minIt && correction < k::NR_itTol) || i >= k::NR_maxIt)
                {
                    const double dT_T2 = std::abs((T_buf - T) / T_buf);
                    timestep           = u;
                    u                  = (dT_T2 == 0 ? k::maxDtStep : k::NR_dTOverTempTarget / dT_T2) * timestep;
                    u                  = std::min(u, timestep * (double)k::maxDtStep);
                    u                  = std::max(u, timestep * (double)k::minDtStep);
                    u                  = std::min(u, (double)k::NR_maxDt);
                    conv               = true;
                }
                else
                    timestep = 0.;
            }
        }
        if (dt_tot - elapsed < dt)
        {
            if (u < dt_tot - elapsed) dt = u;
        }
        else
            dt = u;
        if (tr) tr->push(it, used_dt, elapsed, dt, T_buf, conv ? 1 : 0);

        if (conv)
        {
            for (int j = 0; j < d; ++j)
                Y[j] = Ybuf[j];
            T = T_buf;
            i = 0;
            elapsed += timestep;
            ++st.substeps;
            st.run = 0;
            if ((dt_tot - elapsed) / dt_tot < k::substep_dtTol) return st;
        }
        else if (i == 0)
        {
            ++st.rejections;
            ++st.run;
            if (st.run > st.max_run) st.max_run = st.run;
            // A rejected attempt restarts from the same accepted state; when the new attempt's
            // dt and Newton seed are bit-identical to this one's, it is replayed forever.
            if (same_bits(dt, att_dt) && same_bits(T_buf, att_Tbuf))
            {
                st.status = ST_STALLED;
                return st;
            }
        }
        if (max_iter > 0 && total_iter >= max_iter)
        {
            st.status = ST_MAX_ITER;
            return st;
        }
    }
}

static int dimension_of(int net) { return net == 14 ? 14 : net == 86 ? 86 : 87; }

// util::dynamicBatchSize, nnet_util/algorithm.hpp:79-88
static int dynamic_batch_size(size_t N, int P)
{
    if (N <= (size_t)P) return 1;
    static const float phi        = 1.61803398875;
    int                f          = (int)std::floor(std::log2((float)N / (float)P) / phi);
    int                pow2_f     = 1 << f;
    int                batch_size = (int)std::floor((float)N / ((float)pow2_f * 2.0 * (float)P));
    return std::max(1, batch_size);
}

} // namespace oracle

using namespace oracle;

extern "C"
{
    int oracle_num_reactions(int net) { return reactions_for(net).size(); }

    // Copies the reaction list in the reference's flat layout.
    int oracle_reaction_list(int net, int* reactant_begin, int* product_begin, int* ids, int* nums)
    {
        const Reactions& R = reactions_for(net);
        for (int i = 0; i <= R.size(); ++i)
            reactant_begin[i] = R.reactantBegin[i];
        for (int i = 0; i < R.size(); ++i)
            product_begin[i] = R.productBegin[i];
        for (size_t i = 0; i < R.id.size(); ++i)
            ids[i] = R.id[i], nums[i] = R.num[i];
        return (int)R.id.size();
    }

    void oracle_rates(int net, const double* Y, double T, double rho, double* BE, double* rates, double* drates)
    {
        eval_rates(net, Y, T, rho, BE, rates, drates);
    }

    int oracle_helm_load(const char* path)
    {
        FILE* f = std::fopen(path, "rb");
        if (!f) return -1;
        const size_t cells = (size_t)helm::imax * helm::jmax;
        for (int q = 0; q < 21; ++q)
        {
            g_helm.arr[q].resize(cells);
            if (std::fread(g_helm.arr[q].data(), sizeof(double), cells, f) != cells)
            {
                std::fclose(f);
                return -2;
            }
        }
        std::fclose(f);
        g_helm.finish();
        return 0;
    }

    // out[13] = cv, dpdT, p, cp, c, u, dse, dpe, dsp, cv_gaz, cp_gaz, c_gaz, dudYe
    void oracle_eos(int kind, double mu, const double* Z, int zdim, const double* Y, double T, double rho, double* out)
    {
        EosSpec e{kind, mu, Z, zdim};
        EosOut  r = eval_eos(e, Y, T, rho);
        std::memcpy(out, &r, sizeof(r));
    }

    void oracle_helm_raw(double abar_, double zbar_, double T, double rho, double* out)
    {
        EosOut r = helm::eos(g_helm, abar_, zbar_, T, rho);
        std::memcpy(out, &r, sizeof(r));
    }

    void oracle_solve(double* M, double* RHS, double* X, int n) { gauss_solve(M, RHS, X, n, k::epsilonSystem); }

    // The biquintic interpolation of the loaded table alone (helmholtz.cpp:357-557 for ye = 1):
    // out = free, df_d, df_t, df_tt, df_dt at (din, T).  Used to pin the interpolation.
    void oracle_helm_interp(double din, double temp, double* out)
    {
        const helm::Table& tb = g_helm;
        using namespace helm;
        int jat = int((std::log10(temp) - tlo) * tstpi);
        jat     = std::max<int>(1, std::min<int>(jat, jmax - 2));
        int iat = int((std::log10(din) - dlo) * dstpi);
        iat     = std::max<int>(1, std::min<int>(iat, imax - 2));
        double    fi[36];
        const int ks5[9] = {F, FT, FTT, FD, FDD, FDT, FDDT, FDTT, FDDTT};
        for (int q = 0; q < 9; ++q)
        {
            fi[4 * q + 0] = tb.at(ks5[q], iat + 0, jat + 0);
            fi[4 * q + 1] = tb.at(ks5[q], iat + 1, jat + 0);
            fi[4 * q + 2] = tb.at(ks5[q], iat + 0, jat + 1);
            fi[4 * q + 3] = tb.at(ks5[q], iat + 1, jat + 1);
        }
        const double xt = std::max<double>((temp - tb.t_[jat]) * tb.dti_sav[jat], 0.);
        const double xd = std::max<double>((din - tb.d[iat]) * tb.ddi_sav[iat], 0.);
        const double mxt = 1. - xt, mxd = 1. - xd;
        const double si0t = psi0(xt), si1t = psi1(xt) * tb.dt_sav[jat], si2t = psi2(xt) * tb.dt2_sav[jat];
        const double si0mt = psi0(mxt), si1mt = -psi1(mxt) * tb.dt_sav[jat], si2mt = psi2(mxt) * tb.dt2_sav[jat];
        const double si0d = psi0(xd), si1d = psi1(xd) * tb.dd_sav[iat], si2d = psi2(xd) * tb.dd2_sav[iat];
        const double si0md = psi0(mxd), si1md = -psi1(mxd) * tb.dd_sav[iat], si2md = psi2(mxd) * tb.dd2_sav[iat];
        const double dsi0t = dpsi0(xt) * tb.dti_sav[jat], dsi1t = dpsi1(xt), dsi2t = dpsi2(xt) * tb.dt_sav[jat];
        const double dsi0mt = -dpsi0(mxt) * tb.dti_sav[jat], dsi1mt = dpsi1(mxt), dsi2mt = -dpsi2(mxt) * tb.dt_sav[jat];
        const double dsi0d = dpsi0(xd) * tb.ddi_sav[iat], dsi1d = dpsi1(xd), dsi2d = dpsi2(xd) * tb.dd_sav[iat];
        const double dsi0md = -dpsi0(mxd) * tb.ddi_sav[iat], dsi1md = dpsi1(mxd), dsi2md = -dpsi2(mxd) * tb.dd_sav[iat];
        const double ddsi0t = ddpsi0(xt) * tb.dt2i_sav[jat], ddsi1t = ddpsi1(xt) * tb.dti_sav[jat], ddsi2t = ddpsi2(xt);
        const double ddsi0mt = ddpsi0(mxt) * tb.dt2i_sav[jat], ddsi1mt = -ddpsi1(mxt) * tb.dti_sav[jat],
                     ddsi2mt = ddpsi2(mxt);
        out[0] = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, si0d, si1d, si2d, si0md, si1md, si2md);
        out[1] = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);
        out[2] = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
        out[3] = h5(fi, ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
        out[4] = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);
    }

    // One assembly (prepareSystemNR-style inputs are given explicitly).
    int oracle_prepare(int net, int eos_kind, double mu, const double* Z, int zdim, const double* Y, double T,
                       const double* Yg, double Tg, double rho, double drho_dt, double dt, double* Mp, double* RHS)
    {
        const int        d = dimension_of(net);
        const Reactions& R = reactions_for(net);
        EosSpec          e{eos_kind, mu, Z, zdim};
        EosOut           eos = eval_eos(e, Yg, Tg, rho);
        std::vector<double> rates(R.size());
        return prepare_from_guess(net, R, d, Mp, RHS, rates.data(), Y, T, Yg, Tg, rho, drho_dt, eos, dt) ? 0 : 1;
    }

    // One particle through solveSystemSubstep.  stats[4] = status, substeps, newton, rejections.
    // trace (optional) receives 6 doubles per Newton iteration; returns the iteration count.
    long oracle_particle(int net, int eos_kind, double mu, const double* Z, int zdim, double* Y, double* T, double rho,
                         double drho_dt, double dt_tot, double* dt, long max_iter, double* stats, double* trace,
                         int trace_cap)
    {
        const int d = dimension_of(net);
        Work      w;
        w.size(d, reactions_for(net).size());
        EosSpec       e{eos_kind, mu, Z, zdim};
        Trace         tr;
        tr.cap             = trace_cap;
        tr.rec             = trace;
        ParticleStats st   = solve_particle(net, d, e, w, Y, *T, rho, drho_dt, dt_tot, *dt, max_iter, trace ? &tr : nullptr);
        stats[0]           = st.status;
        stats[1]           = st.substeps;
        stats[2]           = st.newton;
        stats[3]           = st.rejections;
        stats[4]           = st.max_run;
        return tr.len;
    }

    // computeHelmholtzImpl (parallel_nuclear_net_impl.cpp:116-141) over particles [0, n): per
    // particle abar_ = sum_j Y[j][i] (std::accumulate from 0., index order) and
    // zbar_ = sum_j Y[j][i] Z[j] (eigen::dot, eigen.hpp:157-166, index order), then the Helmholtz
    // EOS (helmholtz.cpp:340-965); writes u, cv, p, c and dP/dT.  Y: dim pointers to per-species
    // arrays.  nthreads <= 0 keeps the OpenMP default.
    void oracle_helm_eos_batch(size_t n, int dim, const double* Z, const double* temp, const double* rho,
                               double** Y, double* u, double* cv, double* p, double* c, double* dpdT, int nthreads)
    {
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; ++i)
        {
            double abar = 0.;
            for (int j = 0; j < dim; ++j)
                abar += Y[j][i];
            double zbar = 0.;
            for (int j = 0; j < dim; ++j)
                zbar += Y[j][i] * Z[j];
            const EosOut r = helm::eos(g_helm, abar, zbar, temp[i], rho[i]);
            u[i]           = r.u;
            cv[i]          = r.cv;
            p[i]           = r.p;
            c[i]           = r.c;
            dpdT[i]        = r.dpdT;
        }
    }

    // computeNuclearReactionsImpl (parallel_nuclear_net_impl.cpp:63-103) over particles [0, n).
    // Y is an array of d pointers to per-species arrays (SoA), as in the reference.
    // status / substeps / newton may be null.  nthreads <= 0 keeps the OpenMP default.
    void oracle_batch(int net, int eos_kind, double mu, const double* Z, int zdim, size_t n, double* rho,
                      double* rho_m1, double** Y, double* temp, double* dt, double hydro_dt, double previous_dt,
                      int use_drhodt, uint8_t* status, uint32_t* substeps, uint32_t* newton, int nthreads,
                      long max_iter, uint32_t* max_run, uint8_t* redone)
    {
        if (redone) std::fill(redone, redone + n, (uint8_t)0);
        const int d  = dimension_of(net);
        const int nR = reactions_for(net).size();
        if (nthreads > 0) omp_set_num_threads(nthreads);
        int num_threads = 1;
#pragma omp parallel
#pragma omp master
        num_threads = omp_get_num_threads();
        const int chunk = dynamic_batch_size(n, num_threads);
        EosSpec   e{eos_kind, mu, Z, zdim};

#pragma omp parallel
        {
            Work w;
            w.size(d, nR);
            std::vector<double> Yl(d);
#pragma omp for schedule(dynamic, chunk)
            for (size_t i = 0; i < n; ++i)
            {
                ParticleStats st;
                if (rho[i] > k::minRho && temp[i] > k::minTemp)
                {
                    for (int j = 0; j < d; ++j)
                        Yl[j] = Y[j][i];
                    // use_drhodt 1: parallel_nuclear_net_impl.cpp:90-92, the CPU path overwrites
                    // rho_m1 with drho/dt and integrates with drho_dt = 0.  use_drhodt 2:
                    // parallel_nuclear_net.cu:79-81, drho/dt from rho_m1 (left unchanged) drives the
                    // integration.
                    double drho_dt = 0;
                    if (use_drhodt == 1)
                    {
                        if (rho_m1[i] != 0) rho_m1[i] = (rho[i] - rho_m1[i]) / previous_dt;
                    }
                    else if (use_drhodt == 2)
                    {
                        if (rho_m1[i] != 0) drho_dt = (rho[i] - rho_m1[i]) / previous_dt;
                    }
                    const double T0 = temp[i], dt0 = dt[i];
                    st = solve_particle(net, d, e, w, Yl.data(), temp[i], rho[i], drho_dt, hydro_dt, dt[i], max_iter,
                                        nullptr);
                    if (st.abandoned)
                    {
                        // the device's redo pass: the whole step again in FP64
                        for (int j = 0; j < d; ++j)
                            Yl[j] = Y[j][i];
                        temp[i] = T0, dt[i] = dt0;
                        force_fp64 = true;
                        st = solve_particle(net, d, e, w, Yl.data(), temp[i], rho[i], drho_dt, hydro_dt, dt[i],
                                            max_iter, nullptr);
                        force_fp64 = false;
                        if (redone) redone[i] = 1;
                    }
                    for (int j = 0; j < d; ++j)
                        Y[j][i] = Yl[j];
                }
                else
                    st.status = ST_SKIPPED;
                if (status) status[i] = st.status;
                if (substeps) substeps[i] = st.substeps;
                if (newton) newton[i] = st.newton;
                if (max_run) max_run[i] = st.max_run;
            }
        }
    }

    int oracle_dynamic_batch_size(size_t n, int p) { return dynamic_batch_size(n, p); }

    // Test-only sensitivity probe: 1 moves a deterministic 3/8 of the rate functions'
    // exp/log/pow results by one ulp; 0 restores the exact restatement.
    void oracle_set_ulp_perturbation(int on) { ulp_perturbation() = on; }

    // Linear-solve model of the batch path: 0 the reference's FP64 eigen::solve; 1 FP32 LU with
    // `sweeps` FP64-residual refinement sweeps; 2 GMRES-IR with `sweeps` outer and `inner` GMRES
    // iterations; tol is the correction tolerance.  Counts: solves, converged solves, GMRES
    // iterations (reset by this call).
    void oracle_set_solver_model(int mode, double tol, int sweeps, int inner)
    {
        solver_model()  = SolverModel{mode, tol, sweeps, inner};
        solver_counts() = SolverCounts{};
    }
    void oracle_solver_counts(long* out)
    {
        out[0] = solver_counts().solves;
        out[1] = solver_counts().converged;
        out[2] = solver_counts().inner;
    }
}
