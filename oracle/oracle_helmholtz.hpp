// TEST INFRASTRUCTURE — part of the CPU oracle (see oracle_rates.hpp for the usage rule).
//
// FP64 restatement of the reference Helmholtz EOS, parameterization/eos/helmholtz.cpp:89-965,
// with the table grid of helmholtz.hpp:53-88.  The table itself is supplied at run time
// (oracle_helm_load) in the binary layout written by the package's table generator: 21 arrays
// of [IMAX][JMAX] doubles in the order f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt, dpdf,
// dpdfd, dpdft, dpdfdt, ef, efd, eft, efdt, xf, xfd, xft, xfdt (density index major), which
// holds exactly the values the reference parses from its text table.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace oracle
{
struct EosOut
{
    double cv, dpdT, p;
    double cp, c, u;
    double dse, dpe, dsp;
    double cv_gaz, cp_gaz, c_gaz;
    double dudYe;
};

namespace helm
{
const int    imax = 541, jmax = 201;
const double tlo = 3., thi = 13., tstp = (thi - tlo) / (double)(jmax - 1), tstpi = 1. / tstp;
const double dlo = -12., dhi = 15., dstp = (dhi - dlo) / (double)(imax - 1), dstpi = 1. / dstp;

// helmholtz.hpp:91-126
const double pi      = 3.14159265358979323846264338327950288;
const double h       = 6.6260689633e-27;
const double qe      = 4.8032042712e-10;
const double avo     = 6.0221417930e23;
const double clight  = 2.99792458e10;
const double kerg    = 1.380650424e-16;
const double amu     = 1.66053878283e-24;
const double ssol    = 5.6704e-5;
const double asol    = 4.0 * ssol / clight;
const double kergavo = kerg * avo;
const double sioncon = (2.0 * pi * amu * kerg) / (h * h);
const double a1 = -0.898004, b1 = 0.96786, c1 = 0.220703, d1 = -0.86097, e1 = 2.5269;
const double a2 = 0.29561, b2 = 1.9885, c2 = 0.288675;
const double esqu = qe * qe;

struct Table
{
    bool                ok = false;
    std::vector<double> arr[21]; // each imax*jmax, [i][j]
    double              d[imax], dd_sav[imax - 1], dd2_sav[imax - 1], ddi_sav[imax - 1], dd2i_sav[imax - 1],
        dd3i_sav[imax - 1];
    double t_[jmax], dt_sav[jmax - 1], dt2_sav[jmax - 1], dti_sav[jmax - 1], dt2i_sav[jmax - 1], dt3i_sav[jmax - 1];

    // helmholtz.cpp:101-169 (grid and spacings)
    void finish()
    {
        for (int i = 0; i < imax; ++i)
        {
            double dsav = dlo + i * dstp;
            d[i]        = std::pow((double)10., dsav);
        }
        for (int j = 0; j < jmax; ++j)
        {
            double tsav = tlo + j * tstp;
            t_[j]       = std::pow((double)10., tsav);
        }
        for (int j = 0; j < jmax - 1; ++j)
        {
            const double dth = t_[j + 1] - t_[j], dt2 = dth * dth, dti = 1. / dth, dt2i = 1. / dt2, dt3i = dt2i * dti;
            dt_sav[j]   = dth;
            dt2_sav[j]  = dt2;
            dti_sav[j]  = dti;
            dt2i_sav[j] = dt2i;
            dt3i_sav[j] = dt3i;
        }
        for (int i = 0; i < imax - 1; ++i)
        {
            const double dd = d[i + 1] - d[i], dd2 = dd * dd, ddi = 1. / dd, dd2i = 1. / dd2, dd3i = dd2i * ddi;
            dd_sav[i]   = dd;
            dd2_sav[i]  = dd2;
            ddi_sav[i]  = ddi;
            dd2i_sav[i] = dd2i;
            dd3i_sav[i] = dd3i;
        }
        ok = true;
    }
    double at(int k, int i, int j) const { return arr[k][(size_t)i * jmax + j]; }
};

enum
{
    F = 0, FD, FT, FDD, FTT, FDT, FDDT, FDTT, FDDTT, DPDF, DPDFD, DPDFT, DPDFDT, EF, EFD, EFT, EFDT, XF, XFD, XFT, XFDT
};

// quintic / cubic Hermite basis, helmholtz.cpp:222-311
inline double psi0(double z) { return z * z * z * (z * (-6. * z + 15.) - 10.) + 1.; }
inline double dpsi0(double z) { return z * z * (z * (-30. * z + 60.) - 30.); }
inline double ddpsi0(double z) { return z * (z * (-120. * z + 180.) - 60.); }
inline double psi1(double z) { return z * (z * z * (z * (-3. * z + 8.) - 6.) + 1.); }
inline double dpsi1(double z) { return z * z * (z * (-15. * z + 32.) - 18.) + 1.; }
inline double ddpsi1(double z) { return z * (z * (-60. * z + 96.) - 36.); }
inline double psi2(double z) { return 0.5 * z * z * (z * (z * (-z + 3.) - 3.) + 1.); }
inline double dpsi2(double z) { return 0.5 * z * (z * (z * (-5. * z + 12.) - 9.) + 2.); }
inline double ddpsi2(double z) { return 0.5 * (z * (z * (-20. * z + 36.) - 18.) + 2.); }
inline double xpsi0(double z) { return z * z * (2. * z - 3.) + 1.; }
inline double xdpsi0(double z) { return z * (6. * z - 6.); }
inline double xpsi1(double z) { return z * (z * (z - 2.) + 1.); }
inline double xdpsi1(double z) { return z * (3. * z - 4.) + 1.; }

// helmholtz.cpp:273-286: sum over the 36 products in the reference's term order
inline double h5(const double* fi, double w0t, double w1t, double w2t, double w0mt, double w1mt, double w2mt,
                 double w0d, double w1d, double w2d, double w0md, double w1md, double w2md)
{
    return fi[0] * w0d * w0t + fi[1] * w0md * w0t + fi[2] * w0d * w0mt + fi[3] * w0md * w0mt + fi[4] * w0d * w1t +
           fi[5] * w0md * w1t + fi[6] * w0d * w1mt + fi[7] * w0md * w1mt + fi[8] * w0d * w2t + fi[9] * w0md * w2t +
           fi[10] * w0d * w2mt + fi[11] * w0md * w2mt + fi[12] * w1d * w0t + fi[13] * w1md * w0t + fi[14] * w1d * w0mt +
           fi[15] * w1md * w0mt + fi[16] * w2d * w0t + fi[17] * w2md * w0t + fi[18] * w2d * w0mt +
           fi[19] * w2md * w0mt + fi[20] * w1d * w1t + fi[21] * w1md * w1t + fi[22] * w1d * w1mt +
           fi[23] * w1md * w1mt + fi[24] * w2d * w1t + fi[25] * w2md * w1t + fi[26] * w2d * w1mt +
           fi[27] * w2md * w1mt + fi[28] * w1d * w2t + fi[29] * w1md * w2t + fi[30] * w1d * w2mt +
           fi[31] * w1md * w2mt + fi[32] * w2d * w2t + fi[33] * w2md * w2t + fi[34] * w2d * w2mt + fi[35] * w2md * w2mt;
}
// helmholtz.cpp:314-322
inline double h3(const double* fi, double w0t, double w1t, double w0mt, double w1mt, double w0d, double w1d,
                 double w0md, double w1md)
{
    return fi[0] * w0d * w0t + fi[1] * w0md * w0t + fi[2] * w0d * w0mt + fi[3] * w0md * w0mt + fi[4] * w0d * w1t +
           fi[5] * w0md * w1t + fi[6] * w0d * w1mt + fi[7] * w0md * w1mt + fi[8] * w1d * w0t + fi[9] * w1md * w0t +
           fi[10] * w1d * w0mt + fi[11] * w1md * w0mt + fi[12] * w1d * w1t + fi[13] * w1md * w1t + fi[14] * w1d * w1mt +
           fi[15] * w1md * w1mt;
}

// helmholtz.cpp:340-965
inline EosOut eos(const Table& tb, double abar_, double zbar_, double temp, double rho)
{
    double fi[36];
    double abar = 1 / abar_;
    double zbar = zbar_ / abar_;

    // getTableIndices, helmholtz.cpp:325-337
    int iat, jat;
    {
        const double ye  = std::max((double)1e-16, zbar / abar);
        const double din = ye * rho;
        jat              = int((std::log10(temp) - tlo) * tstpi);
        jat              = std::max<int>(1, std::min<int>(jat, jmax - 2));
        iat              = int((std::log10(din) - dlo) * dstpi);
        iat              = std::max<int>(1, std::min<int>(iat, imax - 2));
    }

    double ytot1 = 1 / abar;
    double ye    = std::max<double>((double)1e-16, zbar / abar);
    double din   = ye * rho;

    double rhoi  = 1. / rho;
    double tempi = 1. / temp;
    double kt    = kerg * temp;
    double ktinv = 1. / kt;

    // radiation
    double prad    = asol * temp * temp * temp * temp / 3;
    double dpraddd = 0.;
    double dpraddt = 4. * prad * tempi;
    double erad    = 3. * prad * rhoi;
    double deraddd = -erad * rhoi;
    double deraddt = 3. * dpraddt * rhoi;
    double srad    = (prad * rhoi + erad) * tempi;
    double dsraddd = (dpraddd * rhoi - prad * rhoi * rhoi + deraddd) * tempi;
    double dsraddt = (dpraddt * rhoi + deraddt - srad) * tempi;
    (void)dsraddt;

    // ions
    double xni     = avo * ytot1 * rho;
    double dxnidd  = avo * ytot1;
    double dxnida  = -xni * ytot1;
    double pion    = xni * kt;
    double dpiondd = dxnidd * kt;
    double dpiondt = xni * kerg;
    double dpionda = dxnida * kt;
    double dpiondz = 0.;
    double eion    = 1.5 * pion * rhoi;
    double deiondd = (1.5 * dpiondd - eion) * rhoi;
    double deiondt = 1.5 * dpiondt * rhoi;
    double deionda = 1.5 * dpionda * rhoi;
    double deiondz = 0.;

    double x = abar * abar * std::sqrt(abar) * rhoi / avo;
    double s = sioncon * temp;
    double z = x * s * std::sqrt(s);
    double y = std::log(z);

    double sion    = (pion * rhoi + eion) * tempi + kergavo * ytot1 * y;
    double dsiondd = (dpiondd * rhoi - pion * rhoi * rhoi + deiondd) * tempi - kergavo * rhoi * ytot1;
    double dsiondt = (dpiondt * rhoi + deiondt) * tempi - (pion * rhoi + eion) * tempi * tempi + 1.5 * kergavo * tempi * ytot1;
    x              = avo * kerg / abar;
    double dsionda = (dpionda * rhoi + deionda) * tempi + kergavo * ytot1 * ytot1 * (2.5 - y);
    double dsiondz = 0.;

    // electron-positron: biquintic Hermite of the free energy, helmholtz.cpp:431-557
    const int ks5[9] = {F, FT, FTT, FD, FDD, FDT, FDDT, FDTT, FDDTT};
    for (int q = 0; q < 9; ++q)
    {
        fi[4 * q + 0] = tb.at(ks5[q], iat + 0, jat + 0);
        fi[4 * q + 1] = tb.at(ks5[q], iat + 1, jat + 0);
        fi[4 * q + 2] = tb.at(ks5[q], iat + 0, jat + 1);
        fi[4 * q + 3] = tb.at(ks5[q], iat + 1, jat + 1);
    }

    double xt  = std::max<double>((temp - tb.t_[jat]) * tb.dti_sav[jat], 0.);
    double xd  = std::max<double>((din - tb.d[iat]) * tb.ddi_sav[iat], 0.);
    double mxt = 1. - xt;
    double mxd = 1. - xd;

    double si0t  = psi0(xt);
    double si1t  = psi1(xt) * tb.dt_sav[jat];
    double si2t  = psi2(xt) * tb.dt2_sav[jat];
    double si0mt = psi0(mxt);
    double si1mt = -psi1(mxt) * tb.dt_sav[jat];
    double si2mt = psi2(mxt) * tb.dt2_sav[jat];
    double si0d  = psi0(xd);
    double si1d  = psi1(xd) * tb.dd_sav[iat];
    double si2d  = psi2(xd) * tb.dd2_sav[iat];
    double si0md = psi0(mxd);
    double si1md = -psi1(mxd) * tb.dd_sav[iat];
    double si2md = psi2(mxd) * tb.dd2_sav[iat];

    double dsi0t  = dpsi0(xt) * tb.dti_sav[jat];
    double dsi1t  = dpsi1(xt);
    double dsi2t  = dpsi2(xt) * tb.dt_sav[jat];
    double dsi0mt = -dpsi0(mxt) * tb.dti_sav[jat];
    double dsi1mt = dpsi1(mxt);
    double dsi2mt = -dpsi2(mxt) * tb.dt_sav[jat];
    double dsi0d  = dpsi0(xd) * tb.ddi_sav[iat];
    double dsi1d  = dpsi1(xd);
    double dsi2d  = dpsi2(xd) * tb.dd_sav[iat];
    double dsi0md = -dpsi0(mxd) * tb.ddi_sav[iat];
    double dsi1md = dpsi1(mxd);
    double dsi2md = -dpsi2(mxd) * tb.dd_sav[iat];

    double ddsi0t  = ddpsi0(xt) * tb.dt2i_sav[jat];
    double ddsi1t  = ddpsi1(xt) * tb.dti_sav[jat];
    double ddsi2t  = ddpsi2(xt);
    double ddsi0mt = ddpsi0(mxt) * tb.dt2i_sav[jat];
    double ddsi1mt = -ddpsi1(mxt) * tb.dti_sav[jat];
    double ddsi2mt = ddpsi2(mxt);

    double free  = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    double df_d  = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);
    double df_t  = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    double df_tt = h5(fi, ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    double df_dt = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);

    // bicubic Hermite weights, helmholtz.cpp:562-585
    si0t   = xpsi0(xt);
    si1t   = xpsi1(xt) * tb.dt_sav[jat];
    si0mt  = xpsi0(mxt);
    si1mt  = -xpsi1(mxt) * tb.dt_sav[jat];
    si0d   = xpsi0(xd);
    si1d   = xpsi1(xd) * tb.dd_sav[iat];
    si0md  = xpsi0(mxd);
    si1md  = -xpsi1(mxd) * tb.dd_sav[iat];
    dsi0t  = xdpsi0(xt) * tb.dti_sav[jat];
    dsi1t  = xdpsi1(xt);
    dsi0mt = -xdpsi0(mxt) * tb.dti_sav[jat];
    dsi1mt = xdpsi1(mxt);
    dsi0d  = xdpsi0(xd) * tb.ddi_sav[iat];
    dsi1d  = xdpsi1(xd);
    dsi0md = -xdpsi0(mxd) * tb.ddi_sav[iat];
    dsi1md = xdpsi1(mxd);

    auto load4 = [&](int k0, int k1, int k2, int k3)
    {
        const int ks[4] = {k0, k1, k2, k3};
        for (int q = 0; q < 4; ++q)
        {
            fi[4 * q + 0] = tb.at(ks[q], iat + 0, jat + 0);
            fi[4 * q + 1] = tb.at(ks[q], iat + 1, jat + 0);
            fi[4 * q + 2] = tb.at(ks[q], iat + 0, jat + 1);
            fi[4 * q + 3] = tb.at(ks[q], iat + 1, jat + 1);
        }
    };

    load4(DPDF, DPDFT, DPDFD, DPDFDT);
    double dpepdd = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);
    dpepdd        = std::max<double>(ye * dpepdd, (double)1.e-30);

    load4(EF, EFT, EFD, EFDT);
    double etaele = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);
    (void)etaele;
    x             = h3(fi, si0t, si1t, si0mt, si1mt, dsi0d, dsi1d, dsi0md, dsi1md);

    load4(XF, XFT, XFD, XFDT);
    double xnefer = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);
    (void)xnefer;
    x             = h3(fi, si0t, si1t, si0mt, si1mt, dsi0d, dsi1d, dsi0md, dsi1md);
    x             = std::max<double>(x, (double)1e-30);

    // electron-positron thermodynamics, helmholtz.cpp:679-703
    x             = din * din;
    double pele   = x * df_d;
    double dpepdt = x * df_dt;
    s             = dpepdd / ye - 2.0 * din * df_d;
    double dpepda = -ytot1 * (2.0 * pele + s * din);
    double dpepdz = rho * ytot1 * (2.0 * din * df_d + s);

    x             = ye * ye;
    double sele   = -df_t * ye;
    double dsepdt = -df_tt * ye;
    double dsepdd = -df_dt * x;
    double dsepda = ytot1 * (ye * df_dt * din - sele);
    double dsepdz = -ytot1 * (ye * df_dt * rho + df_t);

    double eele   = ye * free + temp * sele;
    double deepdt = temp * dsepdt;
    double deepdd = x * df_d + temp * dsepdd;
    double deepda = -ye * ytot1 * (free + df_d * din) + temp * dsepda;
    double deepdz = ytot1 * (free + ye * df_d * rho) + temp * dsepdz;

    // Coulomb corrections, helmholtz.cpp:719-824
    z           = pi * 4. / 3.;
    s           = z * xni;
    double dsdd = z * dxnidd;
    double dsda = z * dxnida;

    double lami     = std::pow((double)1. / s, (double)1. / 3.);
    double inv_lami = 1. / lami;
    z               = -lami / 3;
    double lamidd   = z * dsdd / s;
    double lamida   = z * dsda / s;

    double plasg   = zbar * zbar * esqu * ktinv * inv_lami;
    z              = -plasg * inv_lami;
    double plasgdd = z * lamidd;
    double plasgda = z * lamida;
    double plasgdt = -plasg * ktinv * kerg;
    double plasgdz = 2.0 * plasg / zbar;

    double ecoul = 0, pcoul = 0, scoul = 0, decouldd = 0, decouldt = 0, decoulda = 0, decouldz = 0, dpcouldd = 0,
           dpcouldt = 0, dpcoulda = 0, dpcouldz = 0, dscouldd = 0, dscouldt = 0, dscoulda = 0, dscouldz = 0;

    if (plasg >= 1.)
    {
        x        = std::pow(plasg, (double)0.25);
        y        = avo * ytot1 * kerg;
        ecoul    = y * temp * (a1 * plasg + b1 * x + c1 / x + d1);
        pcoul    = rho * ecoul / 3.;
        scoul    = -y * (3.0 * b1 * x - 5.0 * c1 / x + d1 * (std::log(plasg) - 1.) - e1);
        y        = avo * ytot1 * kt * (a1 + 0.25 / plasg * (b1 * x - c1 / x));
        decouldd = y * plasgdd;
        decouldt = y * plasgdt + ecoul / temp;
        decoulda = y * plasgda - ecoul / abar;
        decouldz = y * plasgdz;
        y        = rho / 3.;
        dpcouldd = ecoul + y * decouldd / 3.;
        dpcouldt = y * decouldt;
        dpcoulda = y * decoulda;
        dpcouldz = y * decouldz;
        y        = -avo * kerg / (abar * plasg) * (0.75 * b1 * x + 1.25 * c1 / x + d1);
        dscouldd = y * plasgdd;
        dscouldt = y * plasgdt;
        dscoulda = y * plasgda - scoul / abar;
        dscouldz = y * plasgdz;
    }
    else if (plasg < 1.)
    {
        x        = plasg * std::sqrt(plasg);
        y        = std::pow(plasg, (double)b2);
        z        = c2 * x - a2 * y / 3.;
        pcoul    = -pion * z;
        ecoul    = 3.0 * pcoul / rho;
        scoul    = -avo / abar * kerg * (c2 * x - a2 * (b2 - 1.) / b2 * y);
        s        = 1.5 * c2 * x / plasg - a2 * b2 * y / plasg / 3.;
        dpcouldd = -dpiondd * z - pion * s * plasgdd;
        dpcouldt = -dpiondt * z - pion * s * plasgdt;
        dpcoulda = -dpionda * z - pion * s * plasgda;
        dpcouldz = -dpiondz * z - pion * s * plasgdz;
        s        = 3.0 / rho;
        decouldd = s * dpcouldd - ecoul / rho;
        decouldt = s * dpcouldt;
        decoulda = s * dpcoulda;
        decouldz = s * dpcouldz;
        s        = -avo * kerg / (abar * plasg) * (1.5 * c2 * x - a2 * (b2 - 1.) * y);
        dscouldd = s * plasgdd;
        dscouldt = s * plasgdt;
        dscoulda = s * plasgda - scoul / abar;
        dscouldz = s * plasgdz;
    }

    x = prad + pion + pele + pcoul;
    y = erad + eion + eele + ecoul;
    z = srad + sion + sele + scoul;
    if (x <= 0. || y <= 0.)
    {
        pcoul = dpcouldd = dpcouldt = dpcoulda = dpcouldz = 0.;
        ecoul = decouldd = decouldt = decoulda = decouldz = 0.;
        scoul = dscouldd = dscouldt = dscoulda = dscouldz = 0.;
    }

    double pgas    = pion + pele + pcoul;
    double egas    = eion + eele + ecoul;
    double dpgasdd = dpiondd + dpepdd + dpcouldd;
    double dpgasdt = dpiondt + dpepdt + dpcouldt;
    double degasdt = deiondt + deepdt + decouldt;
    double degasdz = deiondz + deepdz + decouldz;
    double dsgasdd = dsiondd + dsepdd + dscouldd;
    double dsgasdt = dsiondt + dsepdt + dscouldt;
    (void)dsionda;
    (void)dsiondz;
    (void)deepda;
    (void)dpepda;
    (void)dpepdz;
    (void)dsepda;
    (void)dsepdz;
    (void)deionda;
    (void)degasdz;

    double pres    = prad + pgas;
    double ener    = erad + egas;
    double dpresdd = dpraddd + dpgasdd;
    double dpresdt = dpraddt + dpgasdt;
    double rhoerdd = deraddd + (deiondd + deepdd + decouldd);
    double rhoerdt = deraddt + degasdt;
    double rhotrdd = dsraddd + dsgasdd;
    double rhotrdt = dsraddt + dsgasdt;

    EosOut res;
    double zz       = pgas * rhoi;
    double zzi      = rho / pgas;
    double chit_gas = temp / pgas * dpgasdt;
    double chid_gas = dpgasdd * zzi;
    res.cv_gaz      = degasdt;
    x               = zz * chit_gas / (temp * res.cv_gaz);
    double gam1_gas = chit_gas * x + chid_gas;
    res.cp_gaz      = res.cv_gaz * gam1_gas / chid_gas;
    z               = 1. + (egas + clight * clight) * zzi;
    res.c_gaz       = clight * std::sqrt(gam1_gas / z);

    zz          = pres * rhoi;
    zzi         = rho / pres;
    double chit = temp / pres * dpresdt;
    double chid = dpresdd * zzi;
    res.cv      = rhoerdt;
    x           = zz * chit / (temp * res.cv);
    double gam1 = chit * x + chid;
    res.cp      = res.cv * gam1 / chid;
    z           = 1. + (ener + clight * clight) * zzi;
    res.c       = clight * std::sqrt(gam1 / z);

    x       = rho * rho;
    res.dse = temp * rhotrdt / rhoerdt - 1.;
    res.dpe = (rhoerdd * x + temp * dpresdt) / pres - 1.;
    res.dsp = -rhotrdd * x / dpresdt - 1.;

    res.dpdT  = dpresdt;
    res.dudYe = (deiondz + deepdz + decouldz) * abar;
    res.p     = pres;
    res.u     = ener;
    return res;
}
} // namespace helm
} // namespace oracle
