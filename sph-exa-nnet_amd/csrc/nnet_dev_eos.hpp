// Device Helmholtz EOS (parameterization/eos/helmholtz.cpp:340-965), with the reference's
// expression trees so that the FP64 results round like the reference's.  One template serves two
// callers:
//   helmholtz_eval<false>  the Newton iteration, which consumes only cv and dP/dT
//                          (nuclear_net.hpp:560, 603-608): radiation, ions, the electron-positron
//                          free energy from the biquintic Hermite interpolation, the Coulomb
//                          corrections and the "bomb proof" guard, restricted to the terms that
//                          reach those two outputs;
//   helmholtz_eval<true>   the post-step refresh (computeHelmholtzImpl,
//                          parallel_nuclear_net_impl.cpp:116-141), which also writes the internal
//                          energy u, the pressure p and the sound speed c; c needs dP/drho and so
//                          the bicubic interpolation of the dpdf tables.
// The entropy terms, eta and n_e (the ef/xf tables) reach none of u, cv, p, c, dP/dT and are left
// out.
//
// Table layout in HBM: for every lower-left cell corner (i, j) the 36 doubles the biquintic
// interpolation needs -- f, ft, ftt, fd, fdd, fdt, fddt, fdtt, fddtt at the four corners
// (i,j), (i+1,j), (i,j+1), (i+1,j+1), in the reference's coefficient order
// (helmholtz.cpp:431-466) -- stored contiguously: one EOS call reads one 288-byte block.  A
// second array holds, per cell, the 16 doubles of the bicubic dP/drho interpolation (dpdf,
// dpdft, dpdfd, dpdfdt at the same four corners, helmholtz.cpp:588-606): 128 bytes.
#pragma once
#include <hip/hip_runtime.h>

namespace nnet_amd
{
constexpr int HELM_IMAX = 541, HELM_JMAX = 201, HELM_CELL = 36, HELM_CELL3 = 16;

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
__device__ __forceinline__ double xpsi0(double z) { return z * z * (2. * z - 3.) + 1.; }
__device__ __forceinline__ double xpsi1(double z) { return z * (z * (z - 2.) + 1.); }

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

// The five biquintic sums of helmholtz.cpp:505-557 -- the free energy and its d, t, tt and dt
// derivatives, i.e. h5 with the weight sets (si, si), (si, dsi), (dsi, si), (ddsi, si),
// (dsi, dsi) for (t, d) -- in one pass over the cell's 36 coefficients.  Each sum keeps h5's term
// order and rounding: term k = fi[k] * wd * wt, added left to right.  Weights are indexed
// 0, m0, 1, m1, 2, m2 (w0, w0m, w1, w1m, w2, w2m); coefficient pair q = (2q, 2q+1) shares the
// t weight PT[q] and takes the d weights PD[q], PD[q] + 1.
// Register pressure: left to itself the compiler forms all 180 products before the sums
// (hundreds of live registers, spilled to scratch).  Empty asm statements that take the
// coefficient pair as it is consumed and the five sums after it as operands pin the order: a
// pair's products cannot start before the previous pair's sums exist.  That keeps 5 sums, 30
// weights and two groups of loads live (the next group is issued before the current one is
// consumed).  The asm statements emit no instructions.
__device__ __forceinline__ void h5_five(const double* __restrict__ c, const double (&sd)[6], const double (&dsd)[6],
                                        const double (&st)[6], const double (&dst)[6], const double (&ddst)[6],
                                        double& free, double& df_d, double& df_t, double& df_tt, double& df_dt)
{
    constexpr int PD[18] = {0, 0, 0, 0, 0, 0, 2, 2, 4, 4, 2, 2, 4, 4, 2, 2, 4, 4};
    constexpr int PT[18] = {0, 1, 2, 3, 4, 5, 0, 1, 0, 1, 2, 3, 2, 3, 4, 5, 4, 5};
    constexpr int G      = 3; // coefficient pairs per load group
    double2       cur[G], nxt[G];
#pragma unroll
    for (int j = 0; j < G; ++j)
        cur[j] = *reinterpret_cast<const double2*>(c + 2 * j);
#pragma unroll
    for (int g0 = 0; g0 < 18; g0 += G)
    {
        if (g0 + G < 18)
#pragma unroll
            for (int j = 0; j < G; ++j)
                nxt[j] = *reinterpret_cast<const double2*>(c + 2 * (g0 + G + j));
#pragma unroll
        for (int j = 0; j < G; ++j)
        {
            const int q = g0 + j, d = PD[q], t = PT[q];
            double2   v = cur[j];
            asm volatile("" : "+v"(v.x), "+v"(v.y));
            if (q == 0)
            {
                free  = v.x * sd[d] * st[t];
                df_d  = v.x * dsd[d] * st[t];
                df_t  = v.x * sd[d] * dst[t];
                df_tt = v.x * sd[d] * ddst[t];
                df_dt = v.x * dsd[d] * dst[t];
            }
            else
            {
                free  = free + v.x * sd[d] * st[t];
                df_d  = df_d + v.x * dsd[d] * st[t];
                df_t  = df_t + v.x * sd[d] * dst[t];
                df_tt = df_tt + v.x * sd[d] * ddst[t];
                df_dt = df_dt + v.x * dsd[d] * dst[t];
            }
            free  = free + v.y * sd[d + 1] * st[t];
            df_d  = df_d + v.y * dsd[d + 1] * st[t];
            df_t  = df_t + v.y * sd[d + 1] * dst[t];
            df_tt = df_tt + v.y * sd[d + 1] * ddst[t];
            df_dt = df_dt + v.y * dsd[d + 1] * dst[t];
            asm volatile("" : "+v"(free), "+v"(df_d), "+v"(df_t), "+v"(df_tt), "+v"(df_dt));
        }
#pragma unroll
        for (int j = 0; j < G; ++j)
            cur[j] = nxt[j];
    }
}

// Bicubic Hermite sum in the reference's term order (helmholtz.cpp:314-322).
__device__ __forceinline__ double h3(const double* fi, double w0t, double w1t, double w0mt, double w1mt, double w0d,
                                     double w1d, double w0md, double w1md)
{
    return fi[0] * w0d * w0t + fi[1] * w0md * w0t + fi[2] * w0d * w0mt + fi[3] * w0md * w0mt + fi[4] * w0d * w1t +
           fi[5] * w0md * w1t + fi[6] * w0d * w1mt + fi[7] * w0md * w1mt + fi[8] * w1d * w0t + fi[9] * w1md * w0t +
           fi[10] * w1d * w0mt + fi[11] * w1md * w0mt + fi[12] * w1d * w1t + fi[13] * w1md * w1t + fi[14] * w1d * w1mt +
           fi[15] * w1md * w1mt;
}

// Outputs of the post-step refresh (the fields computeHelmholtzImpl writes).
struct HelmOut
{
    double u, cv, p, c, dpdT;
};

// Helmholtz EOS for abar_ = sum(Y), zbar_ = sum(Y Z) (helmholtz.hpp:217-225,
// helmholtz.cpp:340-955).  FULL = false fills o.cv and o.dpdT only; FULL = true fills every
// field of HelmOut and reads cells3.  The statements shared by both forms are the same, so cv
// and dP/dT are bit-identical between them.
template<bool FULL>
__device__ __forceinline__ void helmholtz_eval(const HelmGrid* __restrict__ g, const double* __restrict__ cells,
                                               const double* __restrict__ cells3, double abar_, double zbar_,
                                               double temp, double rho, HelmOut& o)
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

    // radiation (helmholtz.cpp:369-380)
    const double prad    = asol * temp * temp * temp * temp / 3;
    const double dpraddd = 0.;
    const double dpraddt = 4. * prad * tempi;
    const double erad    = 3. * prad * rhoi;
    const double deraddt = 3. * dpraddt * rhoi;

    // ions (helmholtz.cpp:383-399)
    const double xni     = avo * ytot1 * rho;
    const double dxnidd  = avo * ytot1;
    const double pion    = xni * kt;
    const double dpiondd = dxnidd * kt;
    const double dpiondt = xni * kerg;
    const double eion    = 1.5 * pion * rhoi;
    const double deiondt = 1.5 * dpiondt * rhoi;

    const double xt  = dmax((temp - g->t_[jat]) * g->dti_sav[jat], 0.);
    const double xd  = dmax((din - g->d[iat]) * g->ddi_sav[iat], 0.);
    const double mxt = 1. - xt;
    const double mxd = 1. - xd;

    // quintic Hermite weights (helmholtz.cpp:469-502), in the order 0, m0, 1, m1, 2, m2
    const double st[6]   = {psi0(xt), psi0(mxt), psi1(xt) * g->dt_sav[jat], -psi1(mxt) * g->dt_sav[jat],
                            psi2(xt) * g->dt2_sav[jat], psi2(mxt) * g->dt2_sav[jat]};
    const double sd[6]   = {psi0(xd), psi0(mxd), psi1(xd) * g->dd_sav[iat], -psi1(mxd) * g->dd_sav[iat],
                            psi2(xd) * g->dd2_sav[iat], psi2(mxd) * g->dd2_sav[iat]};
    const double dst[6]  = {dpsi0(xt) * g->dti_sav[jat], -dpsi0(mxt) * g->dti_sav[jat], dpsi1(xt), dpsi1(mxt),
                            dpsi2(xt) * g->dt_sav[jat], -dpsi2(mxt) * g->dt_sav[jat]};
    const double dsd[6]  = {dpsi0(xd) * g->ddi_sav[iat], -dpsi0(mxd) * g->ddi_sav[iat], dpsi1(xd), dpsi1(mxd),
                            dpsi2(xd) * g->dd_sav[iat], -dpsi2(mxd) * g->dd_sav[iat]};
    const double ddst[6] = {ddpsi0(xt) * g->dt2i_sav[jat], ddpsi0(mxt) * g->dt2i_sav[jat],
                            ddpsi1(xt) * g->dti_sav[jat], -ddpsi1(mxt) * g->dti_sav[jat], ddpsi2(xt), ddpsi2(mxt)};

    double free, df_d, df_t, df_tt, df_dt;
    h5_five(cells + ((size_t)iat * HELM_JMAX + jat) * HELM_CELL, sd, dsd, st, dst, ddst, free, df_d, df_t, df_tt,
            df_dt);

    // dP/drho of the electron-positron gas from the bicubic dpdf interpolation
    // (helmholtz.cpp:562-606)
    double dpepdd = 0;
    if constexpr (FULL)
    {
        double        fi3[16];
        const double* c3 = cells3 + ((size_t)iat * HELM_JMAX + jat) * HELM_CELL3;
#pragma unroll
        for (int q = 0; q < 16; q += 2)
        {
            const double2 v = *reinterpret_cast<const double2*>(c3 + q);
            fi3[q]          = v.x;
            fi3[q + 1]      = v.y;
        }
        const double xsi0t  = xpsi0(xt);
        const double xsi1t  = xpsi1(xt) * g->dt_sav[jat];
        const double xsi0mt = xpsi0(mxt);
        const double xsi1mt = -xpsi1(mxt) * g->dt_sav[jat];
        const double xsi0d  = xpsi0(xd);
        const double xsi1d  = xpsi1(xd) * g->dd_sav[iat];
        const double xsi0md = xpsi0(mxd);
        const double xsi1md = -xpsi1(mxd) * g->dd_sav[iat];
        dpepdd              = h3(fi3, xsi0t, xsi1t, xsi0mt, xsi1mt, xsi0d, xsi1d, xsi0md, xsi1md);
        dpepdd              = dmax(ye * dpepdd, 1.e-30);
    }

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
    const double dsdd     = z * dxnidd;
    const double lami     = pow(1. / s, 1. / 3.);
    const double inv_lami = 1. / lami;
    z                     = -lami / 3;
    const double lamidd   = z * dsdd / s;
    const double plasg    = zbar * zbar * esqu * ktinv * inv_lami;
    z                     = -plasg * inv_lami;
    const double plasgdd  = z * lamidd;
    const double plasgdt  = -plasg * ktinv * kerg;

    double ecoul = 0, pcoul = 0, decouldd = 0, decouldt = 0, dpcouldd = 0, dpcouldt = 0;
    if (plasg >= 1.)
    {
        x           = pow(plasg, 0.25);
        double y    = avo * ytot1 * kerg;
        ecoul       = y * temp * (a1 * plasg + b1 * x + c1 / x + d1);
        pcoul       = rho * ecoul / 3.;
        y           = avo * ytot1 * kt * (a1 + 0.25 / plasg * (b1 * x - c1 / x));
        decouldd    = y * plasgdd;
        decouldt    = y * plasgdt + ecoul / temp;
        y           = rho / 3.;
        dpcouldd    = ecoul + y * decouldd / 3.;
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
        dpcouldd = -dpiondd * z - pion * s * plasgdd;
        dpcouldt = -dpiondt * z - pion * s * plasgdt;
        s        = 3.0 / rho;
        decouldd = s * dpcouldd - ecoul / rho;
        decouldt = s * dpcouldt;
    }

    // "bomb proof" guard (helmholtz.cpp:827-850)
    x              = prad + pion + pele + pcoul;
    const double y = erad + eion + eele + ecoul;
    if (x <= 0. || y <= 0.)
    {
        pcoul = dpcouldd = dpcouldt = 0.;
        ecoul = decouldd = decouldt = 0.;
    }

    const double dpgasdt = dpiondt + dpepdt + dpcouldt;
    const double degasdt = deiondt + deepdt + decouldt;
    o.dpdT               = dpraddt + dpgasdt;
    o.cv                 = deraddt + degasdt;
    (void)decouldd;
    if constexpr (FULL)
    {
        // totals and the sound speed (helmholtz.cpp:852-965)
        const double pgas    = pion + pele + pcoul;
        const double egas    = eion + eele + ecoul;
        const double dpgasdd = dpiondd + dpepdd + dpcouldd;
        const double pres    = prad + pgas;
        const double ener    = erad + egas;
        const double dpresdd = dpraddd + dpgasdd;
        const double zz      = pres * rhoi;
        const double zzi     = rho / pres;
        const double chit    = temp / pres * o.dpdT;
        const double chid    = dpresdd * zzi;
        x                    = zz * chit / (temp * o.cv);
        const double gam1    = chit * x + chid;
        z                    = 1. + (ener + clight * clight) * zzi;
        o.c                  = clight * sqrt(gam1 / z);
        o.p                  = pres;
        o.u                  = ener;
    }
}

// cv and dP/dT for the Newton iteration (HelmholtzFunctor, helmholtz.hpp:217-225).
__device__ __forceinline__ void helmholtz_cv_dpdt(const HelmGrid* __restrict__ g, const double* __restrict__ cells,
                                                  double abar_, double zbar_, double temp, double rho, double& cv_out,
                                                  double& dpdT_out)
{
    HelmOut o;
    helmholtz_eval<false>(g, cells, nullptr, abar_, zbar_, temp, rho, o);
    cv_out   = o.cv;
    dpdT_out = o.dpdT;
}

} // namespace nnet_amd
