// Helmholtz free-energy table generator for the electron-positron gas.
//
// Produces the 21 tables the Helmholtz EOS interpolates (grid and order of the reference,
// parameterization/eos/helmholtz.cpp:89-137, helmholtz.hpp:53-88: 541 densities
// log10(rho Ye) in [-12, 15] x 201 temperatures log10 T in [3, 13]) for a gas with Ye = 1:
//   f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt   free energy per gram and its derivatives
//   dpdf, dpdfd, dpdft, dpdfdt                    dP/drho and derivatives
//   ef, efd, eft, efdt                            degeneracy parameter eta and derivatives
//   xf, xfd, xft, xfdt                            e- + e+ number density and derivatives
//
// Physics (Timmes & Arnett 1999, ApJS 125, 277; Timmes & Swesty 2000, ApJS 126, 501):
//   n  = K1 b^{3/2} [F_{1/2}(eta,b) + b F_{3/2}(eta,b)]
//   P  = K2 b^{5/2} [F_{3/2}(eta,b) + (b/2) F_{5/2}(eta,b)]
//   E  = K3 b^{5/2} [F_{3/2}(eta,b) + b F_{5/2}(eta,b)]         (kinetic)
//   F_k(eta, b) = int_0^inf x^k sqrt(1 + b x / 2) / (exp(x - eta) + 1) dx,  b = kT / (m_e c^2)
// positrons with eta+ = -eta - 2/b; eta solves n- - n+ = rho N_A (Ye = 1).  With the pair rest
// mass counted in the energy, the free energy per gram is f = eta k T N_A - P / rho.
//
// Derivatives are exact (up to quadrature error): every quantity is carried as a bivariate
// truncated Taylor series in (u, v) = (drho/rho, dT/T) of degree <= 2 in each variable, and
// eta(rho, T) is obtained by a chord-Newton iteration on the series (each sweep fixes one more
// total degree).  The Fermi-Dirac integrals use fixed composite Gauss-Legendre rules around the
// scalar eta, so the same rule integrates the value and all its derivatives.
//
// Output: binary (21 arrays [541][201] of doubles, the order above) and, optionally, the
// reference's text format: STR( ... ) with, for each temperature j and density i, the nine f
// values, then the dpdf, ef and xf blocks (helmholtz.cpp:106-137).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

namespace
{
const int    IMAX = 541, JMAX = 201;
const double tlo = 3., thi = 13., tstp = (thi - tlo) / (double)(JMAX - 1);
const double dlo = -12., dhi = 15., dstp = (dhi - dlo) / (double)(IMAX - 1);

// physical constants of the reference EOS (helmholtz.hpp:91-115)
const double pi     = 3.14159265358979323846264338327950288;
const double h      = 6.6260689633e-27;
const double avo    = 6.0221417930e23;
const double clight = 2.99792458e10;
const double kerg   = 1.380650424e-16;
const double me     = 9.1093821545e-28;
const double mec2   = me * clight * clight;
const double lam    = h / (me * clight);
const double K1     = 8. * pi * std::sqrt(2.) / (lam * lam * lam);
const double K2     = 16. * pi * std::sqrt(2.) / 3. * mec2 / (lam * lam * lam);
const double K3     = 8. * pi * std::sqrt(2.) * mec2 / (lam * lam * lam);

// ---------------------------------------------------------------------------------------------
// bivariate truncated Taylor series, degree <= 2 in u and in v
// ---------------------------------------------------------------------------------------------
struct S
{
    double c[3][3];
    S() { std::memset(c, 0, sizeof(c)); }
    explicit S(double a)
    {
        std::memset(c, 0, sizeof(c));
        c[0][0] = a;
    }
    double v0() const { return c[0][0]; }
};
inline S operator+(S a, const S& b)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            a.c[i][j] += b.c[i][j];
    return a;
}
inline S operator-(S a, const S& b)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            a.c[i][j] -= b.c[i][j];
    return a;
}
inline S operator*(S a, double s)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            a.c[i][j] *= s;
    return a;
}
inline S operator*(double s, S a) { return a * s; }
inline S operator+(S a, double s)
{
    a.c[0][0] += s;
    return a;
}
inline S operator*(const S& a, const S& b)
{
    S r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double s = 0;
            for (int p = 0; p <= i; ++p)
                for (int q = 0; q <= j; ++q)
                    s += a.c[p][q] * b.c[i - p][j - q];
            r.c[i][j] = s;
        }
    return r;
}
// f(x0 + h) = sum_n d[n] h^n for h without constant term (h^5 = 0)
inline S compose(const S& x, const double d[5])
{
    S hh = x;
    hh.c[0][0] = 0;
    S r(d[0]);
    S p = hh;
    for (int n = 1; n <= 4; ++n)
    {
        r = r + p * d[n];
        if (n < 4) p = p * hh;
    }
    return r;
}
inline S sexp(const S& x)
{
    const double e    = std::exp(x.v0());
    const double d[5] = {e, e, e / 2, e / 6, e / 24};
    return compose(x, d);
}
inline S srecip(const S& x)
{
    const double a = 1. / x.v0();
    const double d[5] = {a, -a * a, a * a * a, -a * a * a * a, a * a * a * a * a};
    return compose(x, d);
}
inline S spow(const S& x, double al)
{
    const double a = x.v0(), p = std::pow(a, al);
    const double d[5] = {p, al * p / a, al * (al - 1) / 2 * p / (a * a), al * (al - 1) * (al - 2) / 6 * p / (a * a * a),
                         al * (al - 1) * (al - 2) * (al - 3) / 24 * p / (a * a * a * a)};
    return compose(x, d);
}

// ---------------------------------------------------------------------------------------------
// Gauss-Legendre rules on [-1, 1]
// ---------------------------------------------------------------------------------------------
struct GL
{
    std::vector<double> x, w;
    explicit GL(int n)
    {
        x.resize(n), w.resize(n);
        for (int i = 0; i < n; ++i)
        {
            double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 0;
            for (int it = 0; it < 100; ++it)
            {
                double p1 = 1, p2 = 0;
                for (int k = 1; k <= n; ++k)
                {
                    const double p3 = p2;
                    p2              = p1;
                    p1              = ((2. * k - 1.) * z * p2 - (k - 1.) * p3) / k;
                }
                pp              = n * (z * p1 - p2) / (z * z - 1.);
                const double z1 = z;
                z               = z1 - p1 / pp;
                if (std::fabs(z - z1) < 1e-16) break;
            }
            x[i] = -z;
            w[i] = 2. / ((1. - z * z) * pp * pp);
        }
    }
};
const GL& gl8()
{
    static GL g(8);
    return g;
}
const GL& gl32()
{
    static GL g(32);
    return g;
}

// Quadrature nodes (x, weight) for F_k(eta, b) with eta near eta0.
struct Nodes
{
    std::vector<double> x, w;
};
void build_nodes(double eta0, Nodes& nd)
{
    nd.x.clear(), nd.w.clear();
    const double a = eta0 > 30 ? eta0 - 30 : 0.; // below a the occupation is 1 to 1e-13
    const double b = (eta0 > 0 ? eta0 : 0.) + 60.;
    auto add_s = [&](double s0, double s1, const GL& g)
    {
        // x = s^2 on s in [s0, s1]
        const double hm = 0.5 * (s1 - s0), c = 0.5 * (s1 + s0);
        for (size_t k = 0; k < g.x.size(); ++k)
        {
            const double s = c + hm * g.x[k];
            nd.x.push_back(s * s);
            nd.w.push_back(g.w[k] * hm * 2. * s);
        }
    };
    auto add_x = [&](double x0, double x1, const GL& g)
    {
        const double hm = 0.5 * (x1 - x0), c = 0.5 * (x1 + x0);
        for (size_t k = 0; k < g.x.size(); ++k)
        {
            nd.x.push_back(c + hm * g.x[k]);
            nd.w.push_back(g.w[k] * hm);
        }
    };
    if (a > 0)
    {
        // smooth part (occupation ~ 1): split in s = sqrt(x) into panels
        const double sa = std::sqrt(a);
        const int    np = 4;
        for (int p = 0; p < np; ++p)
            add_s(sa * p / np, sa * (p + 1) / np, gl32());
        // transition region in x, panels of width 1.5
        const int nx = (int)std::ceil((b - a) / 1.5);
        for (int p = 0; p < nx; ++p)
            add_x(a + (b - a) * p / nx, a + (b - a) * (p + 1) / nx, gl8());
    }
    else
    {
        // x in [0, b] through s = sqrt(x), panels of width 0.125 in s
        const double sb = std::sqrt(b);
        const int    np = (int)std::ceil(sb / 0.125);
        for (int p = 0; p < np; ++p)
            add_s(sb * p / np, sb * (p + 1) / np, gl8());
    }
}

// F_{1/2}, F_{3/2}, F_{5/2} as series for eta (series, scalar part eta0 = node centre) and b.
void fd_integrals(const Nodes& nd, double eta0, const S& eta, const S& b, S F[3])
{
    F[0] = F[1] = F[2] = S();
    S deta = eta;
    deta.c[0][0] -= eta0; // eta = eta0 + deta
    const S em = sexp(deta * -1.); // exp(-(eta - eta0))
    const S ep = sexp(deta);       // exp(eta - eta0)
    for (size_t k = 0; k < nd.x.size(); ++k)
    {
        const double x = nd.x[k];
        const double t = x - eta0;
        if (t > 700) continue;
        // occupation 1 / (exp(x - eta) + 1) without overflow on either side of eta
        S fermi;
        if (t > 0)
        {
            const S w = ep * std::exp(-t); // exp(eta - x) <= ~1
            fermi     = w * srecip(w + 1.);
        }
        else
        {
            const S z = em * std::exp(t); // exp(x - eta) <= ~1
            fermi     = srecip(z + 1.);
        }
        const S sq    = spow(b * (x / 2.) + 1., 0.5);
        const S c     = (sq * fermi) * nd.w[k];
        const double sx = std::sqrt(x);
        F[0]            = F[0] + c * sx;
        F[1]            = F[1] + c * (sx * x);
        F[2]            = F[2] + c * (sx * x * x);
    }
}

// scalar net number density n- - n+ and its eta-derivative
void net_density(double eta, double bt, double& g, double& dg)
{
    Nodes nd;
    double n[2] = {0, 0}, dn[2] = {0, 0};
    for (int sp = 0; sp < 2; ++sp)
    {
        const double e = sp == 0 ? eta : -eta - 2. / bt;
        if (e < -700) continue;
        build_nodes(e, nd);
        double f12 = 0, f32 = 0, d12 = 0, d32 = 0;
        for (size_t k = 0; k < nd.x.size(); ++k)
        {
            const double x = nd.x[k];
            if (x - e > 700) continue;
            const double z   = std::exp(x - e);
            const double fm  = 1. / (z + 1.);
            const double dfm = fm * (z / (z + 1.)); // d fermi / d eta
            const double q   = std::sqrt(1. + bt * x / 2.) * nd.w[k] * std::sqrt(x);
            f12 += q * fm, f32 += q * fm * x, d12 += q * dfm, d32 += q * dfm * x;
        }
        const double pre = K1 * std::pow(bt, 1.5);
        n[sp]            = pre * (f12 + bt * f32);
        dn[sp]           = pre * (d12 + bt * d32) * (sp == 0 ? 1. : -1.);
    }
    g  = n[0] - n[1];
    dg = dn[0] - dn[1];
}

double solve_eta(double target, double bt, double guess)
{
    // bracket
    double lo = guess - 1, hi = guess + 1, g, dg;
    net_density(lo, bt, g, dg);
    while (g > target)
    {
        lo = lo - std::max(1., std::fabs(lo));
        net_density(lo, bt, g, dg);
        if (lo < -1e4) break;
    }
    net_density(hi, bt, g, dg);
    while (g < target)
    {
        hi = hi + std::max(1., std::fabs(hi));
        net_density(hi, bt, g, dg);
        if (hi > 1e12) break;
    }
    double e = std::min(std::max(guess, lo), hi);
    for (int it = 0; it < 200; ++it)
    {
        net_density(e, bt, g, dg);
        const double r = g - target;
        if (r > 0) hi = e;
        else
            lo = e;
        double en = e - r / dg;
        if (!(en > lo && en < hi) || !std::isfinite(en)) en = 0.5 * (lo + hi);
        if (std::fabs(en - e) <= 1e-15 * std::max(1., std::fabs(e)) || hi - lo <= 1e-15 * std::max(1., std::fabs(e)))
        {
            e = en;
            break;
        }
        e = en;
    }
    return e;
}

struct Point
{
    double v[21];
};

// All table quantities at (din, T).
Point evaluate(double din, double temp, double& eta_guess)
{
    const double bt0  = kerg * temp / mec2;
    const double nnet = din * avo;
    const double eta0 = solve_eta(nnet, bt0, eta_guess);
    eta_guess         = eta0;

    S u;
    u.c[1][0] = 1.; // drho / rho
    S v;
    v.c[0][1] = 1.; // dT / T
    const S bt     = (v + 1.) * bt0;
    const S target = (u + 1.) * nnet;

    // chord-Newton on the series: eta <- eta - (n(eta) - target) / (dn/deta at eta0)
    double g0, dg0;
    net_density(eta0, bt0, g0, dg0);
    S     eta(eta0);
    Nodes nde, ndp;
    build_nodes(eta0, nde);
    const double etap0 = -eta0 - 2. / bt0;
    const bool   pos   = etap0 > -700;
    if (pos) build_nodes(etap0, ndp);
    S bt32 = spow(bt, 1.5), bt52 = spow(bt, 2.5);
    S Fe[3], Fp[3];
    for (int it = 0; it < 7; ++it)
    {
        fd_integrals(nde, eta0, eta, bt, Fe);
        S n = K1 * (bt32 * (Fe[0] + bt * Fe[1]));
        if (pos)
        {
            const S etap = eta * -1. - srecip(bt) * 2.;
            fd_integrals(ndp, etap0, etap, bt, Fp);
            n = n - K1 * (bt32 * (Fp[0] + bt * Fp[1]));
        }
        const S r = n - target;
        eta       = eta - r * (1. / dg0);
        eta.c[0][0] = eta0;
    }
    fd_integrals(nde, eta0, eta, bt, Fe);
    S ne = K1 * (bt32 * (Fe[0] + bt * Fe[1]));
    S P  = K2 * (bt52 * (Fe[1] + bt * Fe[2] * 0.5));
    S np, Pp;
    if (pos)
    {
        const S etap = eta * -1. - srecip(bt) * 2.;
        fd_integrals(ndp, etap0, etap, bt, Fp);
        np = K1 * (bt32 * (Fp[0] + bt * Fp[1]));
        Pp = K2 * (bt52 * (Fp[1] + bt * Fp[2] * 0.5));
    }
    const S Ptot = P + Pp;
    const S kT   = (v + 1.) * (kerg * temp);
    // f = eta kT N_A - P / rho  (per gram, Ye = 1)
    const S f  = eta * kT * avo - Ptot * srecip((u + 1.) * din);
    const S xn = ne + np;

    const double r = din, t = temp;
    Point        o;
    o.v[0]  = f.c[0][0];
    o.v[1]  = f.c[1][0] / r;
    o.v[2]  = f.c[0][1] / t;
    o.v[3]  = 2. * f.c[2][0] / (r * r);
    o.v[4]  = 2. * f.c[0][2] / (t * t);
    o.v[5]  = f.c[1][1] / (r * t);
    o.v[6]  = 2. * f.c[2][1] / (r * r * t);
    o.v[7]  = 2. * f.c[1][2] / (r * t * t);
    o.v[8]  = 4. * f.c[2][2] / (r * r * t * t);
    o.v[9]  = Ptot.c[1][0] / r;
    o.v[10] = 2. * Ptot.c[2][0] / (r * r);
    o.v[11] = Ptot.c[1][1] / (r * t);
    o.v[12] = 2. * Ptot.c[2][1] / (r * r * t);
    o.v[13] = eta.c[0][0];
    o.v[14] = eta.c[1][0] / r;
    o.v[15] = eta.c[0][1] / t;
    o.v[16] = eta.c[1][1] / (r * t);
    o.v[17] = xn.c[0][0];
    o.v[18] = xn.c[1][0] / r;
    o.v[19] = xn.c[0][1] / t;
    o.v[20] = xn.c[1][1] / (r * t);
    return o;
}
} // namespace

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s out.bin [out_reference_format.dat]\n", argv[0]);
        return 2;
    }
    std::vector<Point> tab((size_t)IMAX * JMAX);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < JMAX; ++j)
    {
        const double temp = std::pow(10., tlo + j * tstp);
        double       eg   = -5;
        for (int i = 0; i < IMAX; ++i)
        {
            const double din = std::pow(10., dlo + i * dstp);
            tab[(size_t)i * JMAX + j] = evaluate(din, temp, eg);
        }
    }
    for (int i = 0; i < IMAX; ++i)
        for (int j = 0; j < JMAX; ++j)
            for (int q = 0; q < 21; ++q)
                if (!std::isfinite(tab[(size_t)i * JMAX + j].v[q]))
                {
                    std::fprintf(stderr, "non-finite table value: i=%d j=%d q=%d\n", i, j, q);
                    return 1;
                }
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    for (int q = 0; q < 21; ++q)
        for (int i = 0; i < IMAX; ++i)
            for (int j = 0; j < JMAX; ++j)
                std::fwrite(&tab[(size_t)i * JMAX + j].v[q], sizeof(double), 1, f);
    std::fclose(f);
    if (argc > 2)
    {
        FILE* t = std::fopen(argv[2], "w");
        if (!t) return 1;
        std::fprintf(t, "STR(\n");
        for (int j = 0; j < JMAX; ++j)
            for (int i = 0; i < IMAX; ++i)
            {
                const Point& p = tab[(size_t)i * JMAX + j];
                for (int q = 0; q < 9; ++q)
                    std::fprintf(t, "%.17g ", p.v[q]);
                std::fprintf(t, "\n");
            }
        for (int blk = 0; blk < 3; ++blk)
            for (int j = 0; j < JMAX; ++j)
                for (int i = 0; i < IMAX; ++i)
                {
                    const Point& p = tab[(size_t)i * JMAX + j];
                    for (int q = 0; q < 4; ++q)
                        std::fprintf(t, "%.17g ", p.v[9 + 4 * blk + q]);
                    std::fprintf(t, "\n");
                }
        std::fprintf(t, ")\n");
        std::fclose(t);
    }
    return 0;
}
