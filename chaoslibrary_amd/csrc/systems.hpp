// systems.hpp -- per-point step maps for gfx950, one lane per trajectory.
//
// Each system is a compile-time struct with NX (state size), NA (parameter
// count) and a step() that advances X in place, the runtime counterpart of
// the reference's compile-time plug-in (systems/systems.cuh:9, one
// USE_*_MODEL per build).  Kernels are instantiated per system and chosen by
// chaos_system id at launch.  State and parameters live in VGPRs/SGPRs.
//
// Floating point: the library is built with -ffp-contract=off, so every
// product and sum rounds separately unless written as fma(); the fma()
// placements are the reference's __fma_rn calls (systems/systems.cuh:26-58)
// and, for RK4, the canonical order documented in RK4 below.
#pragma once

#include <hip/hip_runtime.h>

namespace chaos {

enum SystemId : int {
    SYS_CHAMELEON = 0,
    SYS_ROSSLER_CD = 1,
    SYS_BASINS = 2,
    SYS_BASINS_2 = 3,
    SYS_DAMIR = 4,
    SYS_ROSSLER_RK4 = 5,
    SYS_LORENZ_RK4 = 6,
    SYS_COUNT = 7
};

// Step constants derived once per launch from h (identical IEEE operations
// on every lane and in the host oracle).
struct StepConst {
    double h;    // step
    double hh;   // 0.5 * h
    double h6;   // h / 6.0
};

__host__ __device__ inline StepConst make_step_const(double h)
{
    StepConst c;
    c.h = h;
    c.hh = 0.5 * h;
    c.h6 = h / 6.0;
    return c;
}

template <int S>
struct System;

// Waves per SIMD requested from the register allocator for a system's
// integration kernels: the RK4 systems fit the 64-VGPR budget of 8 waves
// without spilling; the transcendental maps use the compiler's choice.
// sweep_min is the lower bound given to the sweep kernels: asked for 8
// waves, the allocator caps them at 80 SGPRs and spills the peak pool's
// pointers into VGPR lanes (four v_readlane per append); from 7 it keeps
// them in ~94 SGPRs and still lands at <= 56 VGPRs, i.e. 8 waves.
template <int S>
struct Occupancy {
    static constexpr int waves = 0;
    static constexpr int sweep_min = 1;
};
template <>
struct Occupancy<5> {
    static constexpr int waves = 8;
    static constexpr int sweep_min = 7;
};
template <>
struct Occupancy<6> {
    static constexpr int waves = 8;
    static constexpr int sweep_min = 7;
};

// ---------------------------------------------------------------------------
// Portable sin / cos / exp / atan for the reference systems that call them
// (systems/systems.cuh:28,72-78,93,126-163).  The host and device math
// libraries may round these differently in the last bit, and a chaotic orbit
// turns that into a different orbit (and, where attractors coexist, into a
// different attractor); with these routines a step gives the same bits on
// the GPU and in the CPU oracle, which evaluates the same operations in the
// same order.  Methods (from the mathematics; accuracy within 2 ulp of the C
// library for sin / cos / exp, 4 ulp for atan):
//   sin / cos: x = k pi/2 + r by Cody-Waite reduction with a three-part split
//     of pi/2 (33 + 33 + 53 bits), Taylor series of sin r to r^19 and cos r
//     to r^18 on |r| <= pi/4 (Horner in r^2 with fma), quadrant k mod 4;
//   exp: x = k ln2 + r (two-part split of ln2), Taylor series to r^13, 2^k;
//   atan: |x| > 1 through pi/2 - atan(1/|x|), two half-angle steps
//     atan t = 2 atan(t / (1 + sqrt(1 + t^2))), Taylor series to t^25.
// Coefficients: the series' rationals rounded to double.
// ---------------------------------------------------------------------------
__host__ __device__ inline double pm_trig(double x, int cosine)
{
    if (!(x - x == 0.0)) return x - x;  // inf or NaN -> NaN
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = fma(-k, 0x1.921fb54400000p+0, x);
    r = fma(-k, 0x1.0b4611a600000p-34, r);
    r = fma(-k, 0x1.3198a2e037073p-69, r);
    const double z = r * r;
    double s = -0x1.2f49b46814157p-57;
    s = fma(s, z, 0x1.952c77030ad4ap-49);
    s = fma(s, z, -0x1.ae7f3e733b81fp-41);
    s = fma(s, z, 0x1.6124613a86d09p-33);
    s = fma(s, z, -0x1.ae64567f544e4p-26);
    s = fma(s, z, 0x1.71de3a556c734p-19);
    s = fma(s, z, -0x1.a01a01a01a01ap-13);
    s = fma(s, z, 0x1.1111111111111p-7);
    s = fma(s, z, -0x1.5555555555555p-3);
    double c = -0x1.6827863b97d97p-53;
    c = fma(c, z, 0x1.ae7f3e733b81fp-45);
    c = fma(c, z, -0x1.93974a8c07c9dp-37);
    c = fma(c, z, 0x1.1eed8eff8d898p-29);
    c = fma(c, z, -0x1.27e4fb7789f5cp-22);
    c = fma(c, z, 0x1.a01a01a01a01ap-16);
    c = fma(c, z, -0x1.6c16c16c16c17p-10);
    c = fma(c, z, 0x1.5555555555555p-5);
    c = fma(c, z, -0x1.0000000000000p-1);
    const double sr = fma(r * z, s, r);
    const double cr = fma(z, c, 1.0);
    const int q = ((int)(k - 4.0 * floor(k * 0.25)) + cosine) & 3;  // k mod 4, exact
    return q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
}

__host__ __device__ inline double pm_sin(double x) { return pm_trig(x, 0); }
__host__ __device__ inline double pm_cos(double x) { return pm_trig(x, 1); }

__host__ __device__ inline double pm_exp(double x)
{
    if (x != x) return x;
    if (x > 710.0) return INFINITY;
    if (x < -746.0) return 0.0;
    const double k = rint(x * 0x1.71547652b82fep+0);
    double r = fma(-k, 0x1.62e42fee00000p-1, x);
    r = fma(-k, 0x1.a39ef35793c76p-33, r);
    double p = 0x1.6124613a86d09p-33;
    p = fma(p, r, 0x1.1eed8eff8d898p-29);
    p = fma(p, r, 0x1.ae64567f544e4p-26);
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma(p, r, 0x1.71de3a556c734p-19);
    p = fma(p, r, 0x1.a01a01a01a01ap-16);
    p = fma(p, r, 0x1.a01a01a01a01ap-13);
    p = fma(p, r, 0x1.6c16c16c16c17p-10);
    p = fma(p, r, 0x1.1111111111111p-7);
    p = fma(p, r, 0x1.5555555555555p-5);
    p = fma(p, r, 0x1.5555555555555p-3);
    p = fma(p, r, 0x1.0000000000000p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

__host__ __device__ inline double pm_atan(double x)
{
    if (x != x) return x;
    double a = fabs(x);
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    double t = a / (1.0 + sqrt(fma(a, a, 1.0)));
    t = t / (1.0 + sqrt(fma(t, t, 1.0)));
    const double z = t * t;
    double p = 0x1.47ae147ae147bp-5;
    p = fma(p, z, -0x1.642c8590b2164p-5);
    p = fma(p, z, 0x1.8618618618618p-5);
    p = fma(p, z, -0x1.af286bca1af28p-5);
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5);
    p = fma(p, z, -0x1.1111111111111p-4);
    p = fma(p, z, 0x1.3b13b13b13b14p-4);
    p = fma(p, z, -0x1.745d1745d1746p-4);
    p = fma(p, z, 0x1.c71c71c71c71cp-4);
    p = fma(p, z, -0x1.2492492492492p-3);
    p = fma(p, z, 0x1.999999999999ap-3);
    p = fma(p, z, -0x1.5555555555555p-2);
    double v = 4.0 * fma(t * z, p, t);
    if (inv) v = (0x1.921fb54442d18p+0 - v) + 0x1.1a62633145c07p-54;
    return x < 0 ? -v : v;
}

// systems/systems.cuh:11-39 -- Chameleon, two-stage composition with pm_cos().
template <>
struct System<SYS_CHAMELEON> {
    static constexpr int NX = 3, NA = 7;
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        const double h1 = a[0] * c.h;
        const double h2 = (1 - a[0]) * c.h;
        X[0] = fma(h1, -a[6] * X[1], X[0]);
        X[1] = fma(h1, a[6] * X[0] + a[1] * X[2], X[1]);
        const double cos_term = pm_cos(a[5] * X[1]);
        X[2] = fma(h1, a[2] - a[3] * X[2] + a[4] * cos_term, X[2]);
        X[2] = fma(h2, (a[2] + a[4] * cos_term), X[2]) / (1 + a[3] * h2);
        X[1] = fma(h2, (a[6] * X[0] + a[1] * X[2]), X[1]);
        X[0] = fma(h2, -a[6] * X[1], X[0]);
    }
};

// systems/systems.cuh:41-63 -- Rossler, semi-implicit CD scheme
// (a[0] symmetry parameter, a[1]=a, a[2]=b, a[3]=c).
template <>
struct System<SYS_ROSSLER_CD> {
    static constexpr int NX = 3, NA = 4;
    __host__ __device__ __forceinline__ static void step(double* x, const double* a, const StepConst& c)
    {
        const double h1 = 0.5 * c.h + a[0];
        const double h2 = 0.5 * c.h - a[0];
        x[0] = fma(h1, -x[1] - x[2], x[0]);
        x[1] = fma(h1, x[0] + a[1] * x[1], x[1]);
        x[2] = fma(h1, a[2] + x[2] * (x[0] - a[3]), x[2]);
        double temp = fma(-h2, x[0] - a[3], 1.0);
        x[2] = fma(h2, a[2], x[2]) / temp;
        temp = fma(-h2, a[1], 1.0);
        x[1] = fma(h2, x[0], x[1]) / temp;
        x[0] = fma(h2, -x[1] - x[2], x[0]);
    }
};

// systems/systems.cuh:67-83 -- the reference narrows h2 to float.
template <>
struct System<SYS_BASINS> {
    static constexpr int NX = 3, NA = 5;
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        const double h = c.h;
        const double h2 = (double)(float)(h * (1 - a[0]));
        X[0] = X[0] + h * (pm_sin(X[1]) - a[1] * X[0]);
        X[1] = X[1] + h * (pm_sin(X[2]) - a[1] * X[1]);
        X[2] = X[2] + h * (pm_sin(X[0]) - a[1] * X[2]);
        X[2] = (X[2] + h2 * pm_sin(X[0])) / (1 + h2 * a[1]);
        X[1] = (X[1] + h2 * pm_sin(X[2])) / (1 + h2 * a[1]);
        X[0] = (X[0] + h2 * pm_sin(X[1])) / (1 + h2 * a[1]);
    }
};

// systems/systems.cuh:85-101 -- h1, h2 and the intermediate z are floats.
template <>
struct System<SYS_BASINS_2> {
    static constexpr int NX = 2, NA = 2;
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        const double h1 = (double)(float)(c.h * a[0]);
        const double h2 = (double)(float)(c.h * (1 - a[0]));
        X[0] = X[0] + h1 * (-X[1]);
        X[1] = X[1] + h1 * (a[1] * X[0] + pm_sin(X[1]));
        const double z = (double)(float)X[1];
        X[1] = z + h2 * (a[1] * X[0] + pm_sin(X[1]));
        X[0] = X[0] + h2 * (-X[1]);
    }
};

// systems/systems.cuh:103-178 -- tunnel-diode / memristor circuit, explicit
// Euler, time carried in X[2].
template <>
struct System<SYS_DAMIR> {
    static constexpr int NX = 3, NA = 2;
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        const double h = c.h;
        const double Vc = X[0], XSV = X[1], t = X[2];
        const double U1 = 0.0, U2 = -0.04, C1 = 22e-9;
        const double Is = 1.15E-7, Vt = 1.0 / 15.0, Vp = 0.09, Ip = 2.1e-5;
        const double Iv = -3e-6, D = 26.0, E = 0.14;
        const double Vd = Vc + U2;
        const double Vm = Vc + U1;
        const double idiode = Is * (pm_exp(Vd / Vt) - pm_exp(-Vd / Vt));
        const double itunnel = Ip / Vp * Vd * pm_exp(-(Vd - Vp) / Vp);
        const double iex = Iv * (pm_atan(D * (Vd - E)) + pm_atan(D * (Vd + E)));
        const double Id = idiode + itunnel + iex;
        const double Ron = 1434.0, Roff = 1e6, Von1 = 0.28, Voff1 = 0.14;
        const double Von2 = -0.12, Voff2 = -0.006, TAU = 1e-7, T = 2.0;
        const double boltz = 1.380649e-23, echarge = 1.602176634e-19;
        const double g = XSV / Ron + (1.0 - XSV) / Roff;
        const double arg1 = -1.0 / (T * boltz / echarge) * (Vm - Von1) * (Vm - Von2);
        const double arg2 = -1.0 / (T * boltz / echarge) * (Vm - Voff2) * (Vm - Voff1);
        const double f1 = 1.0 / (1.0 + pm_exp(arg1));
        const double f2 = 1.0 / (1.0 + pm_exp(arg2));
        const double Ix = (1.0 / TAU) * (f1 * (1.0 - XSV) - (1.0 - f2) * XSV);
        const double Imem = Vm * g;
        const double Iin = a[0] * pm_sin(a[1] * t);
        const double dVc = (Iin - Imem - Id) / C1;
        X[0] = Vc + h * dVc;
        X[1] = XSV + h * Ix;
        X[2] = t + h;
    }
};

// Classic RK4 in one canonical operation order (shared with the oracle):
//   k1 = f(X); Y = fma(h/2,k1,X); k2 = f(Y); Y = fma(h/2,k2,X); k3 = f(Y);
//   Y = fma(h,k3,X); k4 = f(Y);
//   s = fma(2,k2,k1); s = fma(2,k3,s); s = s + k4; X = fma(h/6,s,X).
// The weighted sum is accumulated as the stages complete, so only X, Y, k
// and s are live: 12 doubles per trajectory.
template <class F>
__host__ __device__ __forceinline__ void rk4_step(double* X, const double* a, const StepConst& c)
{
    double k[3], Y[3], s[3];
    F::rhs(X, a, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s[i] = k[i];
        Y[i] = fma(c.hh, k[i], X[i]);
    }
    F::rhs(Y, a, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s[i] = fma(2.0, k[i], s[i]);
        Y[i] = fma(c.hh, k[i], X[i]);
    }
    F::rhs(Y, a, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s[i] = fma(2.0, k[i], s[i]);
        Y[i] = fma(c.h, k[i], X[i]);
    }
    F::rhs(Y, a, k);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s[i] = s[i] + k[i];
        X[i] = fma(c.h6, s[i], X[i]);
    }
}

// Rossler x'=-y-z, y'=x+a y, z'=b+z(x-c); a[0]=a, a[1]=b, a[2]=c.
// 4 VALU ops per evaluation (the negation folds into its consumers).
template <>
struct System<SYS_ROSSLER_RK4> {
    static constexpr int NX = 3, NA = 3;
    __host__ __device__ __forceinline__ static void rhs(const double* x, const double* a, double* f)
    {
        f[0] = -(x[1] + x[2]);
        f[1] = fma(a[0], x[1], x[0]);
        f[2] = fma(x[2], x[0] - a[2], a[1]);
    }
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        rk4_step<System>(X, a, c);
    }
};

// Lorenz x'=s(y-x), y'=x(r-z)-y, z'=xy-b z; a[0]=sigma, a[1]=rho, a[2]=beta.
template <>
struct System<SYS_LORENZ_RK4> {
    static constexpr int NX = 3, NA = 3;
    __host__ __device__ __forceinline__ static void rhs(const double* x, const double* a, double* f)
    {
        f[0] = a[0] * (x[1] - x[0]);
        f[1] = fma(x[0], a[1] - x[2], -x[1]);
        f[2] = fma(x[0], x[1], -(a[2] * x[2]));
    }
    __host__ __device__ __forceinline__ static void step(double* X, const double* a, const StepConst& c)
    {
        rk4_step<System>(X, a, c);
    }
};

// Grid value start + k*((end-start)/(n-1)), product and sum rounded
// separately (Bifurcation/src/bifurcationCUDA.cu:400-419,
// LLE/src/LLECUDA.cu:172-203; the host CSV writer uses the same form).
__host__ __device__ inline double grid_value(int k, int n, double start, double end)
{
    if (n == 1) return start;
    const double scale = (end - start) / (n - 1);
    return start + k * scale;
}

// Parameter vector with up to two swept entries, written without dynamic
// register indexing (idxA/idxB are launch-uniform).
template <int NA>
__host__ __device__ __forceinline__ void build_params(double* a, const double* values, int idxA, double va,
                                             int idxB, double vb)
{
#pragma unroll
    for (int k = 0; k < NA; ++k) a[k] = (k == idxA) ? va : ((k == idxB) ? vb : values[k]);
}

// SplitMix64 (Steele, Lea & Flood 2014) -- seeded per point for the
// Benettin perturbation direction.
__host__ __device__ inline unsigned long long splitmix64(unsigned long long& state)
{
    unsigned long long z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uniform direction on the unit sphere by rejection from the cube [-1,1)^NX
// (exact 53-bit uniforms, 1e-6 < r2 <= 1, at most 64 tries), then division
// by sqrt(r2).  Replaces curand XORWOW at LLE/src/LLECUDA.cu:88-101.
template <int NX>
__host__ __device__ inline void unit_vector(unsigned long long seed, long long point, double* v)
{
    unsigned long long st = seed ^ ((unsigned long long)(point + 1) * 0xD1B54A32D192ED03ULL);
    for (int tr = 0; tr < 64; ++tr) {
        double r2 = 0;
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            const unsigned long long z = splitmix64(st);
            v[i] = (double)((long long)(z >> 10) - (1LL << 53)) * 0x1p-53;
        }
#pragma unroll
        for (int i = 0; i < NX; ++i) r2 = r2 + v[i] * v[i];
        if (r2 > 1e-6 && r2 <= 1.0) {
            const double norm = sqrt(r2);
#pragma unroll
            for (int i = 0; i < NX; ++i) v[i] = v[i] / norm;
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) v[i] = i == 0 ? 1.0 : 0.0;
}

}  // namespace chaos
