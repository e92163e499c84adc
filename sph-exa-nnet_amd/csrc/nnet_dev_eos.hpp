// Device EOS evaluations used by the Newton iteration.  The integrator only consumes cv and
// dP/dT (nuclear_net.hpp:560, 603-608), so the device Helmholtz evaluation computes exactly the
// chain that feeds those two outputs (radiation, ions, electron-positron free energy from the
// biquintic Hermite interpolation, Coulomb corrections and the "bomb proof" guard), with the
// reference's expression trees (parameterization/eos/helmholtz.cpp:340-955).  The bicubic
// dP/drho, eta and n_e interpolations and the entropy terms do not reach cv or dP/dT and are
// left out.
//
// Table layout in HBM: for every lower-left cell corner (i, j) the 36 doubles the biquintic
// interpolation needs -- f, ft, ftt, fd, fdd, fdt, fddt, fdtt, fddtt at the four corners
// (i,j), (i+1,j), (i,j+1), (i+1,j+1), in the reference's coefficient order
// (helmholtz.cpp:431-466) -- stored contiguously: one EOS call reads one 288-byte block.
#pragma once
#include <hip/hip_runtime.h>

namespace nnet_amd
{
constexpr int HELM_IMAX = 541, HELM_JMAX = 201, HELM_CELL = 36;

struct HelmGrid
{
    double d[HELM_IMAX], dd_sav[HELM_IMAX], dd2_sav[HELM_IMAX], ddi_sav[HELM_IMAX];
    double t_[HELM_JMAX], dt_sav[HELM_JMAX], dt2_sav[HELM_JMAX], dti_sav[HELM_JMAX], dt2i_sav[HELM_JMAX];
};

namespace helmc
{
constexpr double tlo = 3., thi = 13., tstp = (thi - tlo) / (double)(HELM_JMAX - 1), tstpi = 1. / tstp;
constexpr double dlo = -12., dhi = 15., dstp = (dhi - dlo) / (double)(HELM_IMAX - 1), dstpi = 1. / dstp;
constexpr double pi     = 3.14159265358979323846264338327950288;
constexpr double avo    = 6.0221417930e23;
constexpr double clight = 2.99792458e10;
constexpr double kerg   = 1.380650424e-16;
constexpr double ssol   = 5.6704e-5;
constexpr double asol   = 4.0 * ssol / clight;
constexpr double qe     = 4.8032042712e-10;
constexpr double esqu   = qe * qe;
constexpr double a1 = -0.898004, b1 = 0.96786, c1 = 0.220703, d1 = -0.86097;
constexpr double a2 = 0.29561, b2 = 1.9885, c2 = 0.288675;
} // namespace helmc

__device__ __forceinline__ double psi0(double z) { return z * z * z * (z * (-6. * z + 15.) - 10.) + 1.; }
__device__ __forceinline__ double dpsi0(double z) { return z * z * (z * (-30. * z + 60.) - 30.); }
__device__ __forceinline__ double ddpsi0(double z) { return z * (z * (-120. * z + 180.) - 60.); }
__device__ __forceinline__ double psi1(double z) { return z * (z * z * (z * (-3. * z + 8.) - 6.) + 1.); }
__device__ __forceinline__ double dpsi1(double z) { return z * z * (z * (-15. * z + 32.) - 18.) + 1.; }
__device__ __forceinline__ double ddpsi1(double z) { return z * (z * (-60. * z + 96.) - 36.); }
__device__ __forceinline__ double psi2(double z) { return 0.5 * z * z * (z * (z * (-z + 3.) - 3.) + 1.); }
__device__ __forceinline__ double dpsi2(double z) { return 0.5 * z * (z * (z * (-5. * z + 12.) - 9.) + 2.); }
__device__ __forceinline__ double ddpsi2(double z) { return 0.5 * (z * (z * (-20. * z + 36.) - 18.) + 2.); }

// Biquintic Hermite sum in the reference's term order (helmholtz.cpp:273-286).
__device__ __forceinline__ double h5(const double* fi, double w0t, double w1t, double w2t, double w0mt, double w1mt,
                                     double w2mt, double w0d, double w1d, double w2d, double w0md, double w1md,
                                     double w2md)
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

// cv and dP/dT of the Helmholtz EOS for abar_ = sum(Y), zbar_ = sum(Y Z) (helmholtz.hpp:217-225,
// helmholtz.cpp:340-955).
__device__ __forceinline__ void helmholtz_cv_dpdt(const HelmGrid* __restrict__ g, const double* __restrict__ cells, double abar_,
                                  double zbar_, double temp, double rho, double& cv_out, double& dpdT_out)
{
    using namespace helmc;
    const double abar = 1 / abar_;
    const double zbar = zbar_ / abar_;

    const double ye  = dmax(1e-16, zbar / abar);
    const double din = ye * rho;
    int          jat = int((log10(temp) - tlo) * tstpi);
    jat              = jat < HELM_JMAX - 2 ? jat : HELM_JMAX - 2;
    jat              = jat > 1 ? jat : 1;
    int iat          = int((log10(din) - dlo) * dstpi);
    iat              = iat < HELM_IMAX - 2 ? iat : HELM_IMAX - 2;
    iat              = iat > 1 ? iat : 1;

    const double ytot1 = 1 / abar;
    const double rhoi  = 1. / rho;
    const double tempi = 1. / temp;
    const double kt    = kerg * temp;
    const double ktinv = 1. / kt;

    const double prad    = asol * temp * temp * temp * temp / 3;
    const double dpraddt = 4. * prad * tempi;
    const double erad    = 3. * prad * rhoi;
    const double deraddt = 3. * dpraddt * rhoi;

    const double xni     = avo * ytot1 * rho;
    const double dxnidd  = avo * ytot1;
    const double dxnida  = -xni * ytot1;
    const double pion    = xni * kt;
    const double dpiondd = dxnidd * kt;
    const double dpiondt = xni * kerg;
    const double dpionda = dxnida * kt;
    const double eion    = 1.5 * pion * rhoi;
    const double deiondt = 1.5 * dpiondt * rhoi;
    (void)dpiondd;
    (void)dpionda;

    double fi[36];
    const double* c = cells + ((size_t)iat * HELM_JMAX + jat) * HELM_CELL;
#pragma unroll
    for (int q = 0; q < 36; q += 2)
    {
        const double2 v = *reinterpret_cast<const double2*>(c + q);
        fi[q]           = v.x;
        fi[q + 1]       = v.y;
    }

    const double xt  = dmax((temp - g->t_[jat]) * g->dti_sav[jat], 0.);
    const double xd  = dmax((din - g->d[iat]) * g->ddi_sav[iat], 0.);
    const double mxt = 1. - xt;
    const double mxd = 1. - xd;

    const double si0t  = psi0(xt);
    const double si1t  = psi1(xt) * g->dt_sav[jat];
    const double si2t  = psi2(xt) * g->dt2_sav[jat];
    const double si0mt = psi0(mxt);
    const double si1mt = -psi1(mxt) * g->dt_sav[jat];
    const double si2mt = psi2(mxt) * g->dt2_sav[jat];
    const double si0d  = psi0(xd);
    const double si1d  = psi1(xd) * g->dd_sav[iat];
    const double si2d  = psi2(xd) * g->dd2_sav[iat];
    const double si0md = psi0(mxd);
    const double si1md = -psi1(mxd) * g->dd_sav[iat];
    const double si2md = psi2(mxd) * g->dd2_sav[iat];

    const double dsi0t  = dpsi0(xt) * g->dti_sav[jat];
    const double dsi1t  = dpsi1(xt);
    const double dsi2t  = dpsi2(xt) * g->dt_sav[jat];
    const double dsi0mt = -dpsi0(mxt) * g->dti_sav[jat];
    const double dsi1mt = dpsi1(mxt);
    const double dsi2mt = -dpsi2(mxt) * g->dt_sav[jat];
    const double dsi0d  = dpsi0(xd) * g->ddi_sav[iat];
    const double dsi1d  = dpsi1(xd);
    const double dsi2d  = dpsi2(xd) * g->dd_sav[iat];
    const double dsi0md = -dpsi0(mxd) * g->ddi_sav[iat];
    const double dsi1md = dpsi1(mxd);
    const double dsi2md = -dpsi2(mxd) * g->dd_sav[iat];

    const double ddsi0t  = ddpsi0(xt) * g->dt2i_sav[jat];
    const double ddsi1t  = ddpsi1(xt) * g->dti_sav[jat];
    const double ddsi2t  = ddpsi2(xt);
    const double ddsi0mt = ddpsi0(mxt) * g->dt2i_sav[jat];
    const double ddsi1mt = -ddpsi1(mxt) * g->dti_sav[jat];
    const double ddsi2mt = ddpsi2(mxt);

    const double free  = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    const double df_d  = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);
    const double df_t  = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    const double df_tt = h5(fi, ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
    const double df_dt = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);

    // electron-positron part (helmholtz.cpp:679-703)
    double       x      = din * din;
    const double pele   = x * df_d;
    const double dpepdt = x * df_dt;
    const double sele   = -df_t * ye;
    const double dsepdt = -df_tt * ye;
    const double eele   = ye * free + temp * sele;
    const double deepdt = temp * dsepdt;

    // Coulomb corrections (helmholtz.cpp:719-824)
    double       z        = pi * 4. / 3.;
    double       s        = z * xni;
    const double lami     = pow(1. / s, 1. / 3.);
    const double inv_lami = 1. / lami;
    const double plasg    = zbar * zbar * esqu * ktinv * inv_lami;
    const double plasgdt  = -plasg * ktinv * kerg;

    double ecoul = 0, pcoul = 0, decouldt = 0, dpcouldt = 0;
    if (plasg >= 1.)
    {
        x           = pow(plasg, 0.25);
        double y    = avo * ytot1 * kerg;
        ecoul       = y * temp * (a1 * plasg + b1 * x + c1 / x + d1);
        pcoul       = rho * ecoul / 3.;
        y           = avo * ytot1 * kt * (a1 + 0.25 / plasg * (b1 * x - c1 / x));
        decouldt    = y * plasgdt + ecoul / temp;
        y           = rho / 3.;
        dpcouldt    = y * decouldt;
    }
    else if (plasg < 1.)
    {
        x        = plasg * sqrt(plasg);
        double y = pow(plasg, b2);
        z        = c2 * x - a2 * y / 3.;
        pcoul    = -pion * z;
        ecoul    = 3.0 * pcoul / rho;
        s        = 1.5 * c2 * x / plasg - a2 * b2 * y / plasg / 3.;
        dpcouldt = -dpiondt * z - pion * s * plasgdt;
        s        = 3.0 / rho;
        decouldt = s * dpcouldt;
    }

    // "bomb proof" guard (helmholtz.cpp:827-850)
    x              = prad + pion + pele + pcoul;
    const double y = erad + eion + eele + ecoul;
    if (x <= 0. || y <= 0.)
    {
        decouldt = 0.;
        dpcouldt = 0.;
    }

    const double dpgasdt = dpiondt + dpepdt + dpcouldt;
    const double degasdt = deiondt + deepdt + decouldt;
    dpdT_out             = dpraddt + dpgasdt;
    cv_out               = deraddt + degasdt;
}

} // namespace nnet_amd
