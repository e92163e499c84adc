// Device-side reaction-rate parameterizations for net14 / net86 / net87, evaluated
// cooperatively by the lanes of one workgroup (one particle per workgroup):
//   * one lane per fitted rate (net14: 10, net86/87: 150) computes the forward rate, its
//     reverse from the partition-function ratio, and both T-derivatives;
//   * the six hand-fitted rates (3-alpha, 12C+12C, 12C+16O, 16O+16O, 12C(a,g), 16O(a,g)) with
//     reverses and derivatives: net86 / net87, one group per lane on lanes of different
//     wavefronts (the serial forms, which take the t9 powers RatePow / DerivPow -- functions of
//     T alone -- computed once per wavefront rather than once per group); net14 (one wavefront
//     per particle), their exp / pow values one per lane of the wavefront, then combined (the
//     lane-parallel form);
//   * one lane per species computes the Coulomb-corrected binding energy and the screening
//     chemical potential mu/kT.
// A second phase applies screening per rate and scatters into the reaction-list order.
//
// Every expression keeps the operand order and grouping of the reference so that, with FMA
// contraction disabled, the only differences from the reference CPU build are the last-ulp
// differences between the device exp/log/pow and glibc's.
//   net14: parameterization/net14/net14.cpp:40-660, net14_tables.hpp:47-101
//   net86: parameterization/net86/net86.cpp:41-686
//   net87: parameterization/net87/net87.hpp:117-175, net87.cpp:102-141
#pragma once
#include <hip/hip_runtime.h>

#include "nnet_layout.hpp"

namespace nnet_amd
{
namespace phys
{
constexpr double pi       = 3.14159265358979323846264338327950288;
constexpr double Kb       = 1.380658e-16;
constexpr double Na       = 6.022137e23;
constexpr double e2       = 2.306022645e-19;
constexpr double MevToErg = 9.648529392e17;
} // namespace phys

// std::min / std::max semantics (a NaN first operand is returned unchanged)
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// Rate-bin of the partition-function tables (net14_tables.hpp:71-101).
__device__ __forceinline__ int temperature_bin(double T)
{
    const double e[23] = {1.5e8, 2e8, 3e8, 4e8, 5e8, 6e8, 7e8, 8e8, 9e8, 1e9, 1.5e9, 2e9,
                          2.5e9, 3e9, 3.5e9, 4e9, 4.5e9, 5e9, 6e9, 7e9, 8e9, 9e9, 1e10};
    int k = 0;
#pragma unroll
    for (int q = 0; q < 23; ++q)
        k += (T < e[q]) ? 0 : 1;
    // bins are monotone, so the count of edges not above T is the bin (NaN -> 23 as well)
    return k;
}


// Powers of t9 used by the rate expressions (t9 clamped at 10), net14.cpp:196-213.
struct RatePow
{
    double t9r, t9, t92, t93, t95, t912, t913, t923, t932, t943, t953, t9i, t9i13, t9i23, t9i32, t9rm1, t9r32;
    __device__ explicit RatePow(double T)
    {
        t9r   = T * 1.0e-09;
        t9    = dmin(10., t9r);
        t92   = t9 * t9;
        t93   = t92 * t9;
        t95   = t92 * t93;
        t912  = sqrt(t9);
        t913  = pow(t9, 1. / 3.);
        t923  = t913 * t913;
        t932  = t9 * t912;
        t943  = t9 * t913;
        t953  = t9 * t923;
        t9i   = 1. / t9;
        t9i13 = 1. / t913;
        t9i23 = 1. / t923;
        t9i32 = 1. / t932;
        t9rm1 = 1. / t9r;
        t9r32 = t9 * sqrt(t9r);
    }
};

// Powers of t9 used by the derivative expressions (no clamp), net14.cpp:341-362.
struct DerivPow
{
    double t9, t92, t93, t94, t95, t912, t913, t923, t932, t943, t952, t953, t9i, t9i2, t9i13, t9i23, t9i32, t9i43,
        t9i52, t9i53;
    __device__ explicit DerivPow(double T)
    {
        t9    = T * 1.0e-09;
        t92   = t9 * t9;
        t93   = t92 * t9;
        t94   = t93 * t9;
        t95   = t94 * t9;
        t912  = sqrt(t9);
        t913  = pow(t9, 1. / 3.);
        t923  = t913 * t913;
        t932  = t9 * t912;
        t943  = t9 * t913;
        t952  = t9 * t932;
        t953  = t9 * t923;
        t9i   = 1. / t9;
        t9i2  = t9i * t9i;
        t9i13 = 1. / t913;
        t9i23 = 1. / t923;
        t9i32 = 1. / t932;
        t9i43 = 1. / t943;
        t9i52 = 1. / t952;
        t9i53 = 1. / t953;
    }
};

// Group 0: triple alpha, its reverse, and derivatives (net14.cpp:218-243, 366-423).
__device__ __forceinline__ void special_triple_alpha(double T, const RatePow& p, const DerivPow& q, double& r, double& rl,
                                                     double& dr, double& drl)
{
    {
        const double r2abe = (7.40e+05 * p.t9i32) * exp(-1.0663 * p.t9i) +
                             4.164e+09 * p.t9i23 * exp(-13.49 * p.t9i13 - p.t92 / 0.009604) *
                                 (1. + 0.031 * p.t913 + 8.009 * p.t923 + 1.732 * p.t9 + 49.883 * p.t943 + 27.426 * p.t953);
        const double rbeac = (130. * p.t9i32) * exp(-3.3364 * p.t9i) +
                             2.510e+07 * p.t9i23 * exp(-23.57 * p.t9i13 - p.t92 / 0.055225) *
                                 (1. + 0.018 * p.t913 + 5.249 * p.t923 + 0.650 * p.t9 + 19.176 * p.t943 + 6.034 * p.t953);
        if (T > 8e7)
            r = 2.90e-16 * (r2abe * rbeac) + 0.1 * 1.35e-07 * p.t9i32 * exp(-24.811 * p.t9i);
        else
            r = 2.90e-16 * (r2abe * rbeac) *
                    (0.01 + 0.2 * (1. + 4. * exp(-pow(0.025 * p.t9i, 3.263))) /
                                (1. + 4. * exp(-pow(p.t9 / 0.025, 9.227)))) +
                0.1 * 1.35e-07 * p.t9i32 * exp(-24.811 * p.t9i);
        const double rev = 2.e20 * exp(-84.419412e0 * p.t9i);
        rl               = r * rev * p.t93;
    }
    {
        const double vA  = -24.811 * q.t9i;
        const double vB  = -1.0663 * q.t9i;
        const double vC  = -13.49 * q.t9i13 - q.t92 * 104.123282;
        const double vD  = 1. + .031 * q.t913 + 8.009 * q.t923 + 1.732 * q.t9 + 49.883 * q.t943 + 27.426 * q.t953;
        const double vE  = -3.3364 * q.t9i;
        const double vF  = -23.57 * q.t9i13 - 18.10774106 * q.t92;
        const double vG  = 1. + .018 * q.t913 + 5.249 * q.t923 + .650 * q.t9 + 19.176 * q.t943 + 6.034 * q.t953;
        const double dvA = 24.811 * q.t9i2;
        const double dvB = 1.0663 * q.t9i2;
        const double dvC = q.t9i43 * 13.49 / 3. - 208.246564 * q.t9;
        const double dvD = (q.t9i23 * .031 + q.t9i13 * 16.018 + 5.196 + q.t913 * 199.532 + q.t923 * 137.13) / 3.;
        const double dvE = 3.3364 * q.t9i2;
        const double dvF = q.t9i43 * 23.57 / 3. - 36.21548212 * q.t9;
        const double dvG = (q.t9i23 * .018 + q.t9i13 * 10.498 + 1.950 + q.t913 * 76.704 + q.t923 * 30.17) / 3.;
        const double eB = exp(vB), eC = exp(vC), eE = exp(vE), eF = exp(vF);
        const double r2abe  = 7.4e5 * q.t9i32 * eB + 4.164e9 * q.t9i23 * vD * eC;
        const double rbeac  = 130. * q.t9i32 * eE + 2.510e7 * q.t9i23 * vG * eF;
        const double dr2abe = eB * (-1.11e6 * q.t9i52 + 7.4e5 * q.t9i32 * dvB) +
                              4.164e9 * eC * (-q.t9i53 * vD * 2. / 3. + q.t9i23 * dvC * vD + q.t9i23 * dvD);
        const double drbeac = eE * (-195. * q.t9i52 + 130. * q.t9i32 * dvE) +
                              2.510e7 * eF * (-2. * q.t9i53 * vG / 3. + q.t9i23 * dvF * vG + q.t9i23 * dvG);
        dr = (2.90e-16 * (dr2abe * rbeac + r2abe * drbeac) + 1.35e-8 * exp(vA) * (-1.5 * q.t9i52 + q.t9i32 * dvA)) * 1.e-9;
    }
    {
        const double vA  = -84.419412 * q.t9i;
        const double dvA = 84.419412 * q.t9i2;
        drl              = 2.00e20 * exp(vA) * q.t93 * (dvA * r + 3. * q.t9i * r + dr) * 1.e-9;
    }
}

// Group 1: 12C + 12C (net14.cpp:248-259, 428-443).
__device__ __forceinline__ void special_c12c12(double T, const RatePow& p, const DerivPow& q, double& r, double& dr)
{
    {
        const double  t9a   = p.t9 / (1. + 0.0396 * p.t9);
        const double  t9a13 = pow(t9a, 1. / 3.);
        const double  t9a56 = pow(t9a, 5. / 6.);
        r                   = 4.27e+26 * t9a56 * p.t9i32 * exp(-84.165 / t9a13 - 2.12e-03 * p.t93);
    }
    const double   vA   = q.t9 / (1. + .0396 * q.t9);
    const double   vA56 = pow(vA, 5. / 6.);
    const double   vB   = -84.165 * pow(vA, -1. / 3.) - 2.12e-3 * q.t93;
    const double   dvA  = vA * vA * q.t9i2;
    const double   dvB  = 28.055 * dvA * pow(vA, -4. / 3.) - 6.36e-3 * q.t92;
    dr = 4.27e26 * q.t9i32 * exp(vB) * (pow(vA, -1. / 6.) * dvA * 5. / 6. - 1.5 * vA56 * q.t9i + vA56 * dvB) * 1.e-9;
}

// Group 2: 12C + 16O (net14.cpp:264-281, 448-473).
__device__ __forceinline__ void special_c12o16(double T, const RatePow& p, const DerivPow& q, double& r, double& dr)
{
    r  = 0;
    dr = 0;
    if (T >= 5e8)
    {
        const double  t9ap    = p.t9 / (1. + 0.055 * p.t9);
        const double  t9a2p   = t9ap * t9ap;
        const double  t9a13p  = pow(t9ap, 1. / 3.);
        const double  t9a23p  = t9a13p * t9a13p;
        const double  t9a56ap = pow(t9ap, 5. / 6.);
        r = 1.72e+31 * t9a56ap * p.t9i32 * exp(-106.594 / t9a13p) / (exp(-0.18 * t9a2p) + 1.06e-03 * exp(2.562 * t9a23p));
    }
    const double t9 = T * 1.0e-09;
    if (t9 > .5)
    {
        const double   vA   = q.t9 / (1. + .055 * q.t9);
        const double   vA56 = pow(vA, 5. / 6.);
        const double   vB   = -106.594 * pow(vA, -1. / 3.);
        const double   vC   = -.18 * vA * vA;
        const double   vD   = 2.562 * pow(vA, 2. / 3.);
        const double   eC = exp(vC), eD = exp(vD);
        const double   val  = eC + 1.06e-3 * eD;
        const double   dvA  = vA * vA * q.t9i2;
        const double   dvB  = 106.594 * pow(vA, -4. / 3.) * dvA / 3.;
        const double   dvC  = -.36 * vA * dvA;
        const double   dvD  = 1.708 * dvA * pow(vA, -1. / 3.);
        const double   dval = dvC * eC + 1.06e-3 * dvD * eD;
        dr = 1.72e31 * q.t9i32 * (pow(vA, -1. / 6.) * dvA * 5. / 6. - 1.5 * vA56 * q.t9i + vA56 * (dvB - dval / val)) *
             exp(vB) / val * 1.e-9;
    }
}

// Group 3: 16O + 16O (net14.cpp:286-293, 478-487).
__device__ __forceinline__ void special_o16o16(double T, const RatePow& p, const DerivPow& q, double& r, double& dr)
{
    {
        r = 7.10e+36 * p.t9i23 * exp(-135.93 * p.t9i13 - 0.629 * p.t923 - 0.445 * p.t943 + 0.0103 * p.t92);
    }
    const double   vA  = -135.93e0 * q.t9i13 - .629 * q.t923 - .445 * q.t943 + .0103 * q.t92;
    const double   dvA = 45.31 * q.t9i43 - .629 * q.t9i13 * 2. / 3. - .445 * q.t913 * 4. / 3. + .0206 * q.t9;
    dr                 = 7.10e36 * exp(vA) * q.t9i23 * (-q.t9i * 2. / 3. + dvA) * 1.e-9;
}

// Group 4: 12C(a,g)16O, reverse, derivatives (net14.cpp:298-314, 491-538).
__device__ __forceinline__ void special_c12ag(double T, const RatePow& p, const DerivPow& q, double& r, double& rl, double& dr,
                                              double& drl)
{
    {
        const double  b1 = 1. + 0.0489 * p.t9i23, b2 = 1. + 0.2654 * p.t9i23;
        r = 1.04e+08 / (p.t92 * (b1 * b1)) * exp(-32.120 * p.t9i13 - p.t92 / 12.222016) +
            1.76e+08 / (p.t92 * (b2 * b2)) * exp(-32.120 * p.t9i13) + 1.25e+03 * p.t9i32 * exp(-27.499 * p.t9i) +
            1.43e-02 * p.t95 * exp(-15.541 * p.t9i);
        rl = r * 5.13e+10 * p.t9r32 * exp(-83.108047 * p.t9rm1);
    }
    {
        const double b1  = 1. + .0489 * q.t9i23;
        const double vB  = b1 * b1;
        const double s1  = q.t9 / 3.496;
        const double vC  = -32.120 * q.t9i13 - s1 * s1;
        const double b2  = 1. + .2654 * q.t9i23;
        const double vD  = b2 * b2;
        const double vE  = -32.120 * q.t9i13;
        const double vF  = -27.499 * q.t9i;
        const double vG  = -15.541 * q.t9i;
        const double dvB = -(2. + .0978 * q.t9i23) * (.0326 * q.t9i53);
        const double dvC = 32.120 * q.t9i43 / 3. - 2. * q.t9 / (3.496 * 3.496);
        const double dvD = -(2. + .5308 * q.t9i23) * (.5308 * q.t9i53 / 3.);
        const double dvE = 32.120 * q.t9i43 / 3.;
        const double dvF = 27.499 * q.t9i2;
        const double dvG = 15.541 * q.t9i2;
        dr               = (1.04e8 * exp(vC) * q.t9i2 * (-2. * q.t9i + dvC - dvB / vB) / vB +
              1.76e8 * exp(vE) * q.t9i2 * (-2. * q.t9i + dvE - dvD / vD) / vD +
              1.25e3 * exp(vF) * (-1.5 * q.t9i52 + dvF * q.t9i32) + 1.43e-2 * exp(vG) * (5. * q.t94 + dvG * q.t95)) *
             1.e-9;
    }
    {
        const double vA  = -83.108047 * q.t9i;
        const double dvA = 83.108047 * q.t9i2;
        drl              = 5.13e10 * exp(vA) * (dr * q.t932 + r * 1.5 * q.t912 + r * q.t932 * dvA) * 1.e-9;
    }
}

// Group 5: 16O(a,g)20Ne, reverse, derivatives (net14.cpp:319-334, 519-566).
__device__ __forceinline__ void special_o16ag(double T, const RatePow& p, const DerivPow& q, double& r, double& rl, double& dr,
                                              double& drl)
{
    {
        r  = (9.37e+09 * p.t9i23 * exp(-39.757 * p.t9i13 - p.t92 / 2.515396) + 62.1 * p.t9i32 * exp(-10.297 * p.t9i) +
             538. * p.t9i32 * exp(-12.226 * p.t9i) + 13. * p.t92 * exp(-20.093 * p.t9i));
        rl = r * 5.65e+10 * p.t9r32 * exp(-54.93807 * p.t9rm1);
    }
    {
        const double s1  = q.t9 / 1.586;
        const double vB  = -39.757 * q.t9i13 - s1 * s1;
        const double vC  = -10.297 * q.t9i;
        const double vD  = -12.226 * q.t9i;
        const double vE  = -20.093 * q.t9i;
        const double dvB = 39.757 * q.t9i43 / 3. - 2. * q.t9 / (1.586 * 1.586);
        const double dvC = 10.297 * q.t9i2;
        const double dvD = 12.226 * q.t9i2;
        const double dvE = 20.093 * q.t9i2;
        dr               = (9.37e9 * exp(vB) * (-q.t9i53 * 2. / 3. + q.t9i23 * dvB) +
              62.1 * exp(vC) * (-1.5 * q.t9i52 + q.t9i32 * dvC) + 538. * exp(vD) * (-1.5 * q.t9i52 + q.t9i32 * dvD) +
              13. * exp(vE) * (2. * q.t9 + q.t92 * dvE)) *
             1.e-9;
    }
    {
        const double vA  = -54.903255 * q.t9i;
        const double dvA = 54.903255 * q.t9i2;
        drl              = 5.65e10 * exp(vA) * (dr * q.t932 + 1.5 * r * q.t912 + r * q.t932 * dvA) * 1.e-9;
    }
}

// ---------------------------------------------------------------------------------------------
// The six groups above, lane-parallel over one wavefront (net14, whose one wavefront otherwise
// runs the six serial forms one after another).  Every exp / pow the groups evaluate is a
// function of T only: lane k of the wavefront evaluates value k -- level 1: 36 exps and 13 pows
// of arguments built from the t9 powers; level 2: 6 exps whose arguments need a level-1 pow --
// and every lane then combines the values into the groups' outputs with the serial forms'
// expressions, each exp / pow replaced by its value.  Each value comes from the same argument
// expression through the same function, so the results are bit-identical to the serial forms';
// the wavefront runs one exp pass and one pow pass instead of six chains of them.
// ---------------------------------------------------------------------------------------------
// Level-1 value k < spx::N1: exp(a) or pow(a, b) with the reference's argument expression.
__device__ __forceinline__ double special_level1(int k, const RatePow& p, const DerivPow& q)
{
    double a = 0., b = 1.; // exp(a) for k < NE1, pow(a, b) otherwise
    switch (k)
    {
    // group 0, triple alpha: r2abe, rbeac, r, rev (RatePow), then the derivative exps (DerivPow)
    case 0: a = -1.0663 * p.t9i; break;
    case 1: a = -13.49 * p.t9i13 - p.t92 / 0.009604; break;
    case 2: a = -3.3364 * p.t9i; break;
    case 3: a = -23.57 * p.t9i13 - p.t92 / 0.055225; break;
    case 4: a = -24.811 * p.t9i; break;
    case 5: a = -84.419412e0 * p.t9i; break;
    case 6: a = -1.0663 * q.t9i; break;                           // vB
    case 7: a = -13.49 * q.t9i13 - q.t92 * 104.123282; break;     // vC
    case 8: a = -3.3364 * q.t9i; break;                           // vE
    case 9: a = -23.57 * q.t9i13 - 18.10774106 * q.t92; break;    // vF
    case 10: a = -24.811 * q.t9i; break;                          // vA of dr
    case 11: a = -84.419412 * q.t9i; break;                       // vA of drl
    // group 3, 16O + 16O
    case 12: a = -135.93 * p.t9i13 - 0.629 * p.t923 - 0.445 * p.t943 + 0.0103 * p.t92; break;
    case 13: a = -135.93e0 * q.t9i13 - .629 * q.t923 - .445 * q.t943 + .0103 * q.t92; break;
    // group 4, 12C(a,g)
    case 14: a = -32.120 * p.t9i13 - p.t92 / 12.222016; break;
    case 15: a = -32.120 * p.t9i13; break;
    case 16: a = -27.499 * p.t9i; break;
    case 17: a = -15.541 * p.t9i; break;
    case 18: a = -83.108047 * p.t9rm1; break;
    case 19:
    {
        const double s1 = q.t9 / 3.496;
        a               = -32.120 * q.t9i13 - s1 * s1; // vC
        break;
    }
    case 20: a = -32.120 * q.t9i13; break; // vE
    case 21: a = -27.499 * q.t9i; break;   // vF
    case 22: a = -15.541 * q.t9i; break;   // vG
    case 23: a = -83.108047 * q.t9i; break;
    // group 5, 16O(a,g)
    case 24: a = -39.757 * p.t9i13 - p.t92 / 2.515396; break;
    case 25: a = -10.297 * p.t9i; break;
    case 26: a = -12.226 * p.t9i; break;
    case 27: a = -20.093 * p.t9i; break;
    case 28: a = -54.93807 * p.t9rm1; break;
    case 29:
    {
        const double s1 = q.t9 / 1.586;
        a               = -39.757 * q.t9i13 - s1 * s1; // vB
        break;
    }
    case 30: a = -10.297 * q.t9i; break; // vC
    case 31: a = -12.226 * q.t9i; break; // vD
    case 32: a = -20.093 * q.t9i; break; // vE
    case 33: a = -54.903255 * q.t9i; break;
    // group 2, 12C + 16O: the exps that need no pow
    case 34:
    {
        const double t9ap = p.t9 / (1. + 0.055 * p.t9);
        a                 = -0.18 * (t9ap * t9ap);
        break;
    }
    case 35:
    {
        const double vA = q.t9 / (1. + .055 * q.t9);
        a               = -.18 * vA * vA; // vC
        break;
    }
    // group 1, 12C + 12C: the pows
    case 36: a = p.t9 / (1. + 0.0396 * p.t9), b = 1. / 3.; break;
    case 37: a = p.t9 / (1. + 0.0396 * p.t9), b = 5. / 6.; break;
    case 38: a = q.t9 / (1. + .0396 * q.t9), b = 5. / 6.; break;
    case 39: a = q.t9 / (1. + .0396 * q.t9), b = -1. / 3.; break;
    case 40: a = q.t9 / (1. + .0396 * q.t9), b = -4. / 3.; break;
    case 41: a = q.t9 / (1. + .0396 * q.t9), b = -1. / 6.; break;
    // group 2: the pows
    case 42: a = p.t9 / (1. + 0.055 * p.t9), b = 1. / 3.; break;
    case 43: a = p.t9 / (1. + 0.055 * p.t9), b = 5. / 6.; break;
    case 44: a = q.t9 / (1. + .055 * q.t9), b = 5. / 6.; break;
    case 45: a = q.t9 / (1. + .055 * q.t9), b = -1. / 3.; break;
    case 46: a = q.t9 / (1. + .055 * q.t9), b = 2. / 3.; break;
    case 47: a = q.t9 / (1. + .055 * q.t9), b = -4. / 3.; break;
    case 48: a = q.t9 / (1. + .055 * q.t9), b = -1. / 6.; break;
    default: break;
    }
    return k < spx::NE1 ? exp(a) : pow(a, b);
}

// Level-2 value k < spx::NE2 from the level-1 values v1(j) (wavefront-uniform).
template<class V1>
__device__ __forceinline__ double special_level2(int k, const RatePow& p, const DerivPow& q, const V1& v1)
{
    double a = 0.;
    switch (k)
    {
    case 0: a = -84.165 / v1(36) - 2.12e-03 * p.t93; break;               // c12c12 r: t9a13 = pow(t9a, 1/3)
    case 1: a = -84.165 * v1(39) - 2.12e-3 * q.t93; break;                // c12c12 vB
    case 2: a = -106.594 / v1(42); break;                                // c12o16 r
    case 3: a = 2.562 * (v1(42) * v1(42)); break;                        // c12o16 r: t9a23p
    case 4: a = 2.562 * v1(46); break;                                   // c12o16 vD
    case 5: a = -106.594 * v1(45); break;                                // c12o16 vB
    default: break;
    }
    return exp(a);
}

// The six groups' outputs from the level-1 and level-2 values (the reference's expressions,
// each exp / pow replaced by its value), handed to out(group, r, rl, dr, drl) group by group
// (rl, drl only for groups 0, 4, 5; zero otherwise).
template<class V1, class V2, class Out>
__device__ __forceinline__ void special_combine(double T, const RatePow& p, const DerivPow& q, const V1& v1,
                                                const V2& v2, const Out& out)
{
    // group 0: triple alpha
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        const double r2abe = (7.40e+05 * p.t9i32) * v1(0) +
                             4.164e+09 * p.t9i23 * v1(1) *
                                 (1. + 0.031 * p.t913 + 8.009 * p.t923 + 1.732 * p.t9 + 49.883 * p.t943 + 27.426 * p.t953);
        const double rbeac = (130. * p.t9i32) * v1(2) +
                             2.510e+07 * p.t9i23 * v1(3) *
                                 (1. + 0.018 * p.t913 + 5.249 * p.t923 + 0.650 * p.t9 + 19.176 * p.t943 + 6.034 * p.t953);
        double r;
        if (T > 8e7)
            r = 2.90e-16 * (r2abe * rbeac) + 0.1 * 1.35e-07 * p.t9i32 * v1(4);
        else
            r = 2.90e-16 * (r2abe * rbeac) *
                    (0.01 + 0.2 * (1. + 4. * exp(-pow(0.025 * p.t9i, 3.263))) /
                                (1. + 4. * exp(-pow(p.t9 / 0.025, 9.227)))) +
                0.1 * 1.35e-07 * p.t9i32 * v1(4);
        const double rev = 2.e20 * v1(5);
        gr           = r;
        grl          = r * rev * p.t93;
        const double vD  = 1. + .031 * q.t913 + 8.009 * q.t923 + 1.732 * q.t9 + 49.883 * q.t943 + 27.426 * q.t953;
        const double vG  = 1. + .018 * q.t913 + 5.249 * q.t923 + .650 * q.t9 + 19.176 * q.t943 + 6.034 * q.t953;
        const double dvA = 24.811 * q.t9i2;
        const double dvB = 1.0663 * q.t9i2;
        const double dvC = q.t9i43 * 13.49 / 3. - 208.246564 * q.t9;
        const double dvD = (q.t9i23 * .031 + q.t9i13 * 16.018 + 5.196 + q.t913 * 199.532 + q.t923 * 137.13) / 3.;
        const double dvE = 3.3364 * q.t9i2;
        const double dvF = q.t9i43 * 23.57 / 3. - 36.21548212 * q.t9;
        const double dvG = (q.t9i23 * .018 + q.t9i13 * 10.498 + 1.950 + q.t913 * 76.704 + q.t923 * 30.17) / 3.;
        const double eB = v1(6), eC = v1(7), eE = v1(8), eF = v1(9);
        const double r2abe_ = 7.4e5 * q.t9i32 * eB + 4.164e9 * q.t9i23 * vD * eC;
        const double rbeac_ = 130. * q.t9i32 * eE + 2.510e7 * q.t9i23 * vG * eF;
        const double dr2abe = eB * (-1.11e6 * q.t9i52 + 7.4e5 * q.t9i32 * dvB) +
                              4.164e9 * eC * (-q.t9i53 * vD * 2. / 3. + q.t9i23 * dvC * vD + q.t9i23 * dvD);
        const double drbeac = eE * (-195. * q.t9i52 + 130. * q.t9i32 * dvE) +
                              2.510e7 * eF * (-2. * q.t9i53 * vG / 3. + q.t9i23 * dvF * vG + q.t9i23 * dvG);
        const double dr = (2.90e-16 * (dr2abe * rbeac_ + r2abe_ * drbeac) + 1.35e-8 * v1(10) * (-1.5 * q.t9i52 + q.t9i32 * dvA)) *
                          1.e-9;
        gdr           = dr;
        const double dvA2 = 84.419412 * q.t9i2;
        gdrl          = 2.00e20 * v1(11) * q.t93 * (dvA2 * r + 3. * q.t9i * r + dr) * 1.e-9;
        out(0, gr, grl, gdr, gdrl);
    }
    // group 1: 12C + 12C
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        const double t9a56 = v1(37);
        gr             = 4.27e+26 * t9a56 * p.t9i32 * v2(0);
        const double vA    = q.t9 / (1. + .0396 * q.t9);
        const double vA56  = v1(38);
        const double dvA   = vA * vA * q.t9i2;
        const double dvB   = 28.055 * dvA * v1(40) - 6.36e-3 * q.t92;
        gdr = 4.27e26 * q.t9i32 * v2(1) * (v1(41) * dvA * 5. / 6. - 1.5 * vA56 * q.t9i + vA56 * dvB) * 1.e-9;
        out(1, gr, grl, gdr, gdrl);
    }
    // group 2: 12C + 16O
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        if (T >= 5e8)
        {
            const double t9a56ap = v1(43);
            gr = 1.72e+31 * t9a56ap * p.t9i32 * v2(2) / (v1(34) + 1.06e-03 * v2(3));
        }
        const double t9 = T * 1.0e-09;
        if (t9 > .5)
        {
            const double vA   = q.t9 / (1. + .055 * q.t9);
            const double vA56 = v1(44);
            const double eC = v1(35), eD = v2(4);
            const double val  = eC + 1.06e-3 * eD;
            const double dvA  = vA * vA * q.t9i2;
            const double dvB  = 106.594 * v1(47) * dvA / 3.;
            const double dvC  = -.36 * vA * dvA;
            const double dvD  = 1.708 * dvA * v1(45);
            const double dval = dvC * eC + 1.06e-3 * dvD * eD;
            gdr = 1.72e31 * q.t9i32 * (v1(48) * dvA * 5. / 6. - 1.5 * vA56 * q.t9i + vA56 * (dvB - dval / val)) *
                      v2(5) / val * 1.e-9;
        }
        out(2, gr, grl, gdr, gdrl);
    }
    // group 3: 16O + 16O
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        gr           = 7.10e+36 * p.t9i23 * v1(12);
        const double dvA = 45.31 * q.t9i43 - .629 * q.t9i13 * 2. / 3. - .445 * q.t913 * 4. / 3. + .0206 * q.t9;
        gdr          = 7.10e36 * v1(13) * q.t9i23 * (-q.t9i * 2. / 3. + dvA) * 1.e-9;
        out(3, gr, grl, gdr, gdrl);
    }
    // group 4: 12C(a,g)16O
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        const double b1 = 1. + 0.0489 * p.t9i23, b2 = 1. + 0.2654 * p.t9i23;
        const double r  = 1.04e+08 / (p.t92 * (b1 * b1)) * v1(14) + 1.76e+08 / (p.t92 * (b2 * b2)) * v1(15) +
                         1.25e+03 * p.t9i32 * v1(16) + 1.43e-02 * p.t95 * v1(17);
        gr           = r;
        grl          = r * 5.13e+10 * p.t9r32 * v1(18);
        const double qb1 = 1. + .0489 * q.t9i23;
        const double vB  = qb1 * qb1;
        const double qb2 = 1. + .2654 * q.t9i23;
        const double vD  = qb2 * qb2;
        const double dvB = -(2. + .0978 * q.t9i23) * (.0326 * q.t9i53);
        const double dvC = 32.120 * q.t9i43 / 3. - 2. * q.t9 / (3.496 * 3.496);
        const double dvD = -(2. + .5308 * q.t9i23) * (.5308 * q.t9i53 / 3.);
        const double dvE = 32.120 * q.t9i43 / 3.;
        const double dvF = 27.499 * q.t9i2;
        const double dvG = 15.541 * q.t9i2;
        const double dr  = (1.04e8 * v1(19) * q.t9i2 * (-2. * q.t9i + dvC - dvB / vB) / vB +
                           1.76e8 * v1(20) * q.t9i2 * (-2. * q.t9i + dvE - dvD / vD) / vD +
                           1.25e3 * v1(21) * (-1.5 * q.t9i52 + dvF * q.t9i32) + 1.43e-2 * v1(22) * (5. * q.t94 + dvG * q.t95)) *
                          1.e-9;
        gdr          = dr;
        const double dvA = 83.108047 * q.t9i2;
        gdrl         = 5.13e10 * v1(23) * (dr * q.t932 + r * 1.5 * q.t912 + r * q.t932 * dvA) * 1.e-9;
        out(4, gr, grl, gdr, gdrl);
    }
    // group 5: 16O(a,g)20Ne
    {
        double gr = 0, grl = 0, gdr = 0, gdrl = 0;
        const double r = (9.37e+09 * p.t9i23 * v1(24) + 62.1 * p.t9i32 * v1(25) + 538. * p.t9i32 * v1(26) +
                          13. * p.t92 * v1(27));
        gr           = r;
        grl          = r * 5.65e+10 * p.t9r32 * v1(28);
        const double dvB = 39.757 * q.t9i43 / 3. - 2. * q.t9 / (1.586 * 1.586);
        const double dvC = 10.297 * q.t9i2;
        const double dvD = 12.226 * q.t9i2;
        const double dvE = 20.093 * q.t9i2;
        const double dr  = (9.37e9 * v1(29) * (-q.t9i53 * 2. / 3. + q.t9i23 * dvB) +
                           62.1 * v1(30) * (-1.5 * q.t9i52 + q.t9i32 * dvC) + 538. * v1(31) * (-1.5 * q.t9i52 + q.t9i32 * dvD) +
                           13. * v1(32) * (2. * q.t9 + q.t92 * dvE)) *
                          1.e-9;
        gdr          = dr;
        const double dvA = 54.903255 * q.t9i2;
        gdrl         = 5.65e10 * v1(33) * (dr * q.t932 + 1.5 * r * q.t912 + r * q.t932 * dvA) * 1.e-9;
        out(5, gr, grl, gdr, gdrl);
    }
}

// Coulomb correction functions (net14_tables.hpp:47-67).
__device__ __forceinline__ double coulomb_strong(double x)
{
    const double a = -.898004, b = .96786, c = .220703, d = -.86097;
    const double r = sqrt(sqrt(x));
    return a * x + b * r + c / r + d;
}
__device__ __forceinline__ double coulomb_weak(double x)
{
    const double a = -.5 * 1.7320508075688772 /* sqrt(3.), correctly rounded */, b = .29561, c = 1.9885;
    return a * x * sqrt(x) + b * pow(x, c);
}

// The electron-capture table interpolation of net87 (net87.cpp:102-141), by the 64 lanes of one
// wavefront (k = lane).  The reference finds the grid positions by linear searches (the first
// point not below log T, resp. log(rho Ye)); both grids ascend strictly, so that position is the
// number of grid points below the value, counted here with one ballot per grid instead of up to
// 92 dependent loads.  Lane c < 12 then interpolates rate c with the reference's expression.
__device__ __forceinline__ void electron_interpolate_wave(const double* logT, const double* logR, const double* tab,
                                                          double temp, double rhoElec, double* rate, int k)
{
    constexpr int nT = 41, nR = 51, nC = 12;
    static_assert(nT <= 64 && nR <= 64 && nC <= 64, "one grid point per lane");
    const double lt = log10(temp);
    const double lr = log10(rhoElec);
    const double gT = logT[k < nT ? k : 0];
    const double gR = logR[k < nR ? k : 0];
    int          it = (int)__popcll(__ballot(k < nT && gT < lt));
    int          ir = (int)__popcll(__ballot(k < nR && gR < lr));
    if (k >= nC) return;
    const int it0 = (it - 1 > 0) ? it - 1 : 0, ir0 = (ir - 1 > 0) ? ir - 1 : 0;
    it            = (it < nT - 1) ? it : nT - 1;
    ir            = (ir < nR - 1) ? ir : nR - 1;
    const double x2x  = logT[it] - lt;
    const double xx1  = -logT[it0] + lt;
    const double y2y  = logR[ir] - lr;
    const double yy1  = -logR[ir0] + lr;
    double       x2x1 = logT[it] - logT[it0];
    x2x1              = x2x1 == 0 ? 2 : x2x1;
    double y2y1       = logR[ir] - logR[ir0];
    y2y1              = y2y1 == 0 ? 2 : y2y1;
    const int c       = k;
    rate[c] = (tab[(it0 * nR + ir0) * nC + c] * x2x * y2y + tab[(it * nR + ir0) * nC + c] * xx1 * y2y +
               tab[(it0 * nR + ir) * nC + c] * x2x * yy1 + tab[(it * nR + ir) * nC + c] * xx1 * yy1) /
              (x2x1 * y2y1);
}

} // namespace nnet_amd
