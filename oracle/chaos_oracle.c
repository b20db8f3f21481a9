/*
 * chaos_oracle.c -- CPU restatement of the ChaosLibrary per-point hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Compiled with -ffp-contract=off: every fused multiply-add below is an
 * explicit fma(), every other product and sum rounds separately, exactly as
 * the reference's per-point code compiled for the host with contraction off.
 * Citations are reference paths (file:line).
 */
#include "chaos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_NX 12

/* ------------------------------------------------------------------------ */
/* Portable transcendentals                                                  */
/* ------------------------------------------------------------------------ */

/*
 * sin, cos, exp and atan for the reference systems that call them
 * (systems/systems.cuh:28,72-78,93,126-163).  The kernels evaluate the same
 * operations in the same order (chaoslibrary_amd/csrc/systems.hpp), so a
 * step of these systems gives the same bits on the host and the device,
 * where the platforms' own libraries may differ in the last bit (and chaotic
 * orbits amplify that).  Methods, all from the mathematics:
 *   sin / cos: Cody-Waite reduction x = k pi/2 + r with a three-part split
 *     of pi/2 (33 + 33 + 53 bits, each k * part exact for |k| < 2^20), then
 *     the Taylor series of sin r (to r^19) and cos r (to r^18) on
 *     |r| <= pi/4, Horner in r^2 with fma, quadrant by k mod 4;
 *   exp: x = k ln2 + r (two-part split of ln2), Taylor series to r^13 on
 *     |r| <= ln2/2, scaled by 2^k;
 *   atan: |x| > 1 through pi/2 - atan(1/|x|), two half-angle steps
 *     atan t = 2 atan(t / (1 + sqrt(1 + t^2))) to |t| <= tan(pi/16), then the
 *     Taylor series to t^25.
 * Coefficients are the series' rationals rounded to double.  Pinned by:
 * tests/test_oracle_math.py (within 2 ulp of the C library for sin / cos /
 * exp, 4 ulp for atan, over the arguments the systems meet).
 */
static double pm_trig(double x, int cosine)
{
    if (!(x - x == 0.0)) return x - x;                   /* inf or NaN -> NaN */
    const double k = rint(x * 0x1.45f306dc9c883p-1);     /* nearest multiple of pi/2 */
    double r = fma(-k, 0x1.921fb54400000p+0, x);
    r = fma(-k, 0x1.0b4611a600000p-34, r);
    r = fma(-k, 0x1.3198a2e037073p-69, r);
    const double z = r * r;
    double s = -0x1.2f49b46814157p-57;                    /* -1/19! */
    s = fma(s, z, 0x1.952c77030ad4ap-49);                 /*  1/17! */
    s = fma(s, z, -0x1.ae7f3e733b81fp-41);                /* -1/15! */
    s = fma(s, z, 0x1.6124613a86d09p-33);                 /*  1/13! */
    s = fma(s, z, -0x1.ae64567f544e4p-26);                /* -1/11! */
    s = fma(s, z, 0x1.71de3a556c734p-19);                 /*  1/9!  */
    s = fma(s, z, -0x1.a01a01a01a01ap-13);                /* -1/7!  */
    s = fma(s, z, 0x1.1111111111111p-7);                  /*  1/5!  */
    s = fma(s, z, -0x1.5555555555555p-3);                 /* -1/3!  */
    double c = -0x1.6827863b97d97p-53;                    /* -1/18! */
    c = fma(c, z, 0x1.ae7f3e733b81fp-45);                 /*  1/16! */
    c = fma(c, z, -0x1.93974a8c07c9dp-37);                /* -1/14! */
    c = fma(c, z, 0x1.1eed8eff8d898p-29);                 /*  1/12! */
    c = fma(c, z, -0x1.27e4fb7789f5cp-22);                /* -1/10! */
    c = fma(c, z, 0x1.a01a01a01a01ap-16);                 /*  1/8!  */
    c = fma(c, z, -0x1.6c16c16c16c17p-10);                /* -1/6!  */
    c = fma(c, z, 0x1.5555555555555p-5);                  /*  1/4!  */
    c = fma(c, z, -0x1.0000000000000p-1);                 /* -1/2!  */
    const double sr = fma(r * z, s, r);
    const double cr = fma(z, c, 1.0);
    const int q = ((int)(k - 4.0 * floor(k * 0.25)) + cosine) & 3;  /* k mod 4, exact */
    return q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
}

double orc_sin(double x) { return pm_trig(x, 0); }
double orc_cos(double x) { return pm_trig(x, 1); }

double orc_exp(double x)
{
    if (x != x) return x;
    if (x > 710.0) return INFINITY;
    if (x < -746.0) return 0.0;
    const double k = rint(x * 0x1.71547652b82fep+0);     /* nearest multiple of ln2 */
    double r = fma(-k, 0x1.62e42fee00000p-1, x);
    r = fma(-k, 0x1.a39ef35793c76p-33, r);
    double p = 0x1.6124613a86d09p-33;                     /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);                 /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);                 /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);                 /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);                 /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);                 /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);                 /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);                 /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);                  /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);                  /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);                  /* 1/3!  */
    p = fma(p, r, 0x1.0000000000000p-1);                  /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

double orc_atan(double x)
{
    if (x != x) return x;
    double a = fabs(x);
    const int inv = a > 1.0;
    if (inv) a = 1.0 / a;
    double t = a / (1.0 + sqrt(fma(a, a, 1.0)));
    t = t / (1.0 + sqrt(fma(t, t, 1.0)));
    const double z = t * t;
    double p = 0x1.47ae147ae147bp-5;                      /*  1/25 */
    p = fma(p, z, -0x1.642c8590b2164p-5);                 /* -1/23 */
    p = fma(p, z, 0x1.8618618618618p-5);                  /*  1/21 */
    p = fma(p, z, -0x1.af286bca1af28p-5);                 /* -1/19 */
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5);                  /*  1/17 */
    p = fma(p, z, -0x1.1111111111111p-4);                 /* -1/15 */
    p = fma(p, z, 0x1.3b13b13b13b14p-4);                  /*  1/13 */
    p = fma(p, z, -0x1.745d1745d1746p-4);                 /* -1/11 */
    p = fma(p, z, 0x1.c71c71c71c71cp-4);                  /*  1/9  */
    p = fma(p, z, -0x1.2492492492492p-3);                 /* -1/7  */
    p = fma(p, z, 0x1.999999999999ap-3);                  /*  1/5  */
    p = fma(p, z, -0x1.5555555555555p-2);                 /* -1/3  */
    double v = 4.0 * fma(t * z, p, t);
    if (inv) v = (0x1.921fb54442d18p+0 - v) + 0x1.1a62633145c07p-54;  /* pi/2 - v */
    return x < 0 ? -v : v;
}

/* ------------------------------------------------------------------------ */
/* Systems (L1 plug-in)                                                      */
/* ------------------------------------------------------------------------ */

int orc_dims(int sys, int* nx, int* na)
{
    static const int NX[ORC_SYS_COUNT] = {3, 3, 3, 2, 3, 3, 3};
    static const int NA[ORC_SYS_COUNT] = {7, 4, 5, 2, 2, 3, 3};
    if (sys < 0 || sys >= ORC_SYS_COUNT) return -1;
    if (nx) *nx = NX[sys];
    if (na) *na = NA[sys];
    return 0;
}

/* systems/systems.cuh:11-39 (USE_CHAMELEON_MODEL); __fma_rn -> fma.
 * Pinned by: second-order convergence to its ODE (a[0] = 0.5) against scipy. */
static void step_chameleon(double* X, const double* a, double h)
{
    double h1 = a[0] * h;
    double h2 = (1 - a[0]) * h;
    X[0] = fma(h1, -a[6] * X[1], X[0]);
    X[1] = fma(h1, a[6] * X[0] + a[1] * X[2], X[1]);
    double cos_term = orc_cos(a[5] * X[1]);
    X[2] = fma(h1, a[2] - a[3] * X[2] + a[4] * cos_term, X[2]);
    X[2] = fma(h2, (a[2] + a[4] * cos_term), X[2]) / (1 + a[3] * h2);
    X[1] = fma(h2, (a[6] * X[0] + a[1] * X[2]), X[1]);
    X[0] = fma(h2, -a[6] * X[1], X[0]);
}

/* systems/systems.cuh:41-63 (USE_ROSSLER_MODEL, semi-implicit CD scheme).
 * Pinned by: second-order convergence to the Rossler ODE (a[0] = 0). */
static void step_rossler_cd(double* x, const double* a, double h)
{
    double h1 = 0.5 * h + a[0];
    double h2 = 0.5 * h - a[0];
    x[0] = fma(h1, -x[1] - x[2], x[0]);
    x[1] = fma(h1, x[0] + a[1] * x[1], x[1]);
    x[2] = fma(h1, a[2] + x[2] * (x[0] - a[3]), x[2]);
    double temp = fma(-h2, x[0] - a[3], 1.0);
    x[2] = fma(h2, a[2], x[2]) / temp;
    temp = fma(-h2, a[1], 1.0);
    x[1] = fma(h2, x[0], x[1]) / temp;
    x[0] = fma(h2, -x[1] - x[2], x[0]);
}

/* systems/systems.cuh:67-83 (USE_SYSTEM_FOR_BASINS); h2 is a float there.
 * Pinned by: convergence to its ODE (order 2 for a[0] = 0, order 1 otherwise)
 * against scipy in tests/test_oracle_systems.py. */
static void step_basins(double* X, const double* a, double h)
{
    float h2 = (float)(h * (1 - a[0]));
    X[0] = X[0] + h * (orc_sin(X[1]) - a[1] * X[0]);
    X[1] = X[1] + h * (orc_sin(X[2]) - a[1] * X[1]);
    X[2] = X[2] + h * (orc_sin(X[0]) - a[1] * X[2]);
    X[2] = (X[2] + (double)h2 * orc_sin(X[0])) / (1 + (double)h2 * a[1]);
    X[1] = (X[1] + (double)h2 * orc_sin(X[2])) / (1 + (double)h2 * a[1]);
    X[0] = (X[0] + (double)h2 * orc_sin(X[1])) / (1 + (double)h2 * a[1]);
}

/* systems/systems.cuh:85-101 (USE_SYSTEM_FOR_BASINS_2); h1, h2, z are floats.
 * Pinned by: first-order convergence to its ODE against scipy. */
static void step_basins2(double* X, const double* a, double h)
{
    float h1 = (float)(h * a[0]);
    float h2 = (float)(h * (1 - a[0]));
    X[0] = X[0] + (double)h1 * (-X[1]);
    X[1] = X[1] + (double)h1 * (a[1] * X[0] + orc_sin(X[1]));
    float z = (float)X[1];
    X[1] = (double)z + (double)h2 * (a[1] * X[0] + orc_sin(X[1]));
    X[0] = X[0] + (double)h2 * (-X[1]);
}

/* systems/systems.cuh:103-178 (USE_DAMIR_SYSTEM, explicit Euler, time in X[2]).
 * Pinned by: first-order convergence of both states to the circuit ODE
 * (scipy Radau) and the exact time channel. */
static void step_damir(double* X, const double* a, double h)
{
    double Vc = X[0], XSV = X[1], t = X[2];
    double U1 = 0.0, U2 = -0.04, C1 = 22e-9;
    double Is = 1.15E-7, Vt = 1.0 / 15.0, Vp = 0.09, Ip = 2.1e-5;
    double Iv = -3e-6, D = 26.0, E = 0.14;
    double Vd = Vc + U2;
    double Vm = Vc + U1;
    double idiode = Is * (orc_exp(Vd / Vt) - orc_exp(-Vd / Vt));
    double itunnel = Ip / Vp * Vd * orc_exp(-(Vd - Vp) / Vp);
    double iex = Iv * (orc_atan(D * (Vd - E)) + orc_atan(D * (Vd + E)));
    double Id = idiode + itunnel + iex;
    double Ron = 1434.0, Roff = 1e6, Von1 = 0.28, Voff1 = 0.14;
    double Von2 = -0.12, Voff2 = -0.006, TAU = 1e-7, T = 2.0;
    double boltz = 1.380649e-23, echarge = 1.602176634e-19;
    double g = XSV / Ron + (1.0 - XSV) / Roff;
    double arg1 = -1.0 / (T * boltz / echarge) * (Vm - Von1) * (Vm - Von2);
    double arg2 = -1.0 / (T * boltz / echarge) * (Vm - Voff2) * (Vm - Voff1);
    double exp1 = orc_exp(arg1);
    double exp2 = orc_exp(arg2);
    double f1 = 1.0 / (1.0 + exp1);
    double f2 = 1.0 / (1.0 + exp2);
    double Ix = (1.0 / TAU) * (f1 * (1.0 - XSV) - (1.0 - f2) * XSV);
    double Imem = Vm * g;
    double Iin = a[0] * orc_sin(a[1] * t);
    double dVc = (Iin - Imem - Id) / C1;
    double dXSV = Ix;
    X[0] = Vc + h * dVc;
    X[1] = XSV + h * dXSV;
    X[2] = t + h;
}

/*
 * Classic RK4 (new systems; no reference counterpart, systems/systems.cuh
 * keeps RK4 only as commented templates, debri/cudaLibrary.cu:2927-2975).
 * Canonical operation order shared with the HIP kernels:
 *   k1 = f(X); Y = fma(h/2,k1,X); k2 = f(Y); Y = fma(h/2,k2,X); k3 = f(Y);
 *   Y = fma(h,k3,X); k4 = f(Y);
 *   s = fma(2,k2,k1); s = fma(2,k3,s); s = s + k4; X = fma(h/6,s,X)
 * with h/2 = 0.5*h and h/6 = h/6.0 (IEEE division).
 * Pinned by: closed-form linear growth factor (generic RK4 in the test,
 * not this code), order-4 error decay against scipy DOP853, Lorenz origin
 * fixed point and (x,y,z)->(-x,-y,z) symmetry (bit-exact).
 */
static inline void f_rossler(const double* x, const double* a, double* f)
{
    f[0] = -(x[1] + x[2]);
    f[1] = fma(a[0], x[1], x[0]);
    f[2] = fma(x[2], x[0] - a[2], a[1]);
}

static inline void f_lorenz(const double* x, const double* a, double* f)
{
    f[0] = a[0] * (x[1] - x[0]);
    f[1] = fma(x[0], a[1] - x[2], -x[1]);
    f[2] = fma(x[0], x[1], -(a[2] * x[2]));
}

static void step_rk4(void (*f)(const double*, const double*, double*),
                     double* X, const double* a, double h)
{
    double hh = 0.5 * h, h6 = h / 6.0;
    double k[3], Y[3], s[3];
    int i;
    f(X, a, k);
    for (i = 0; i < 3; ++i) { s[i] = k[i]; Y[i] = fma(hh, k[i], X[i]); }
    f(Y, a, k);
    for (i = 0; i < 3; ++i) { s[i] = fma(2.0, k[i], s[i]); Y[i] = fma(hh, k[i], X[i]); }
    f(Y, a, k);
    for (i = 0; i < 3; ++i) { s[i] = fma(2.0, k[i], s[i]); Y[i] = fma(h, k[i], X[i]); }
    f(Y, a, k);
    for (i = 0; i < 3; ++i) { s[i] = s[i] + k[i]; X[i] = fma(h6, s[i], X[i]); }
}

void orc_step(int sys, double* X, const double* a, double h)
{
    switch (sys) {
    case ORC_SYS_CHAMELEON:   step_chameleon(X, a, h); break;
    case ORC_SYS_ROSSLER_CD:  step_rossler_cd(X, a, h); break;
    case ORC_SYS_BASINS:      step_basins(X, a, h); break;
    case ORC_SYS_BASINS_2:    step_basins2(X, a, h); break;
    case ORC_SYS_DAMIR:       step_damir(X, a, h); break;
    case ORC_SYS_ROSSLER_RK4: step_rk4(f_rossler, X, a, h); break;
    case ORC_SYS_LORENZ_RK4:  step_rk4(f_lorenz, X, a, h); break;
    default: break;
    }
}

void orc_steps(int sys, double* X, const double* a, double h, long long n)
{
    for (long long i = 0; i < n; ++i) orc_step(sys, X, a, h);
}

void orc_integrate(int sys, int n, const double* params, double* states,
                   long long nsteps, double h)
{
    int nx, na;
    if (orc_dims(sys, &nx, &na)) return;
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < n; ++p)
        orc_steps(sys, states + (size_t)p * nx, params + (size_t)p * na, h, nsteps);
}

/* ------------------------------------------------------------------------ */
/* Grid values                                                               */
/* ------------------------------------------------------------------------ */

/* Bifurcation/src/bifurcationCUDA.cu:400-419 and LLE/src/LLECUDA.cu:172-203:
 * start + k * ((end - start) / (n - 1)), product and sum rounded separately. */
double orc_grid_value(int k, int n, double start, double end)
{
    if (n == 1) return start;
    double scale = (end - start) / (n - 1);
    return start + k * scale;
}

/* Basins/src/basinsCUDA.cu:185-189 (pow form, same value as above). */
double orc_grid_value_pow(int idx, int n, double start, double end, int dim)
{
    int k = ((int)((int)idx / pow((double)n, (double)dim))) % n;
    return start + (k * ((double)(end - start) / (double)(n - 1)));
}

/* ------------------------------------------------------------------------ */
/* Integration loop with divergence / fixed-point flags                     */
/* ------------------------------------------------------------------------ */

static int diverged(double v, double maxValue, int check)
{
    if (check == ORC_CHK_BIF_REF) {
        /* Bifurcation/src/bifurcationCUDA.cu:445-453 */
        if (v != v || v == v + 1.0) return 1;
        if (maxValue != 0 && (double)fabsf((float)v) > maxValue) return 1;
    } else if (check == ORC_CHK_BASINS_REF) {
        /* Basins/src/basinsCUDA.cu:350-361 */
        if (isnan(v) || isinf(v)) return 1;
        if (maxValue != 0 && (double)fabsf((float)v) > maxValue) return 1;
    } else {
        if (!isfinite(v)) return 1;
        if (maxValue != 0 && fabs(v) > maxValue) return 1;
    }
    return 0;
}

int orc_diverged(double v, double maxValue, int check) { return diverged(v, maxValue, check); }

static int fixed_point(double t, int check)
{
    if (check == ORC_CHK_BIF_REF)           /* bifurcationCUDA.cu:462 */
        return (double)sqrtf(fabsf((float)t)) < 1e-9;
    if (check == ORC_CHK_BASINS_REF)        /* basinsCUDA.cu:372 */
        return sqrt(fabs(t)) < 1e-9;
    return sqrt(t) < 1e-9;
}

/*
 * Bifurcation/src/bifurcationCUDA.cu:421-467 (preScaller == 1) and
 * Basins/src/basinsCUDA.cu:318-380 (preScaller sub-steps per sample).
 * Stores X[writableVar] before each sample's steps; returns 0 (diverged),
 * -1 (last sample moved less than 1e-9) or 1.  amount <= 0 returns 1 (the
 * reference reads an uninitialised xPrev there).
 * Pinned by: hand-built divergence and equilibrium cases in tests.
 */
int orc_loop(int sys, double* X, const double* a, double h, int amount,
             int preScaller, int writableVar, double maxValue, int check,
             double* data)
{
    int nx, na, i, j;
    double xPrev[ORC_MAX_NX];
    if (orc_dims(sys, &nx, &na)) return 0;
    if (amount <= 0) return 1;
    if (preScaller < 1) preScaller = 1;
    for (i = 0; i < amount; ++i) {
        for (j = 0; j < nx; ++j) xPrev[j] = X[j];
        if (data) data[i] = X[writableVar];
        for (j = 0; j < preScaller; ++j) orc_step(sys, X, a, h);
        if (diverged(X[writableVar], maxValue, check)) return 0;
    }
    double t = 0;
    for (j = 0; j < nx; ++j) t += ((X[j] - xPrev[j]) * (X[j] - xPrev[j]));
    return fixed_point(t, check) ? -1 : 1;
}

/* ------------------------------------------------------------------------ */
/* Peak finder                                                               */
/* ------------------------------------------------------------------------ */

/*
 * Bifurcation/src/bifurcationCUDA.cu:469-514 (= Basins/src/basinsCUDA.cu:382-440),
 * with startDataIndex = 0 and out-of-place outputs (the reference's in-place
 * writes never overtake its reads).  Plateau-aware maxima with 1e-13
 * thresholds; the first peak is dropped and intervals are (t[k+1]-t[k])*h.
 * Pinned by: sampled sines, plateaus, staircases and the rejected-plateau
 * skip case, worked by hand in tests/test_oracle_peaks.py.
 */
int orc_peak_finder(const double* data, int n, double* outPeaks,
                    double* timeOfPeaks, double h)
{
    int amountOfPeaks = 0;
    for (int i = 1; i < n - 1; ++i) {
        if (data[i] - data[i - 1] > 1e-13 && data[i] >= data[i + 1]) {
            for (int j = i; j < n - 1; ++j) {
                if (data[j] < data[j + 1]) {
                    i = j + 1;
                    break;
                }
                if (data[j] - data[j + 1] > 1e-13) {
                    if (outPeaks) outPeaks[amountOfPeaks] = data[j];
                    if (timeOfPeaks)
                        timeOfPeaks[amountOfPeaks] = trunc(((double)j + (double)i) / (double)2);
                    ++amountOfPeaks;
                    i = j + 1;
                    break;
                }
            }
        }
    }
    if (amountOfPeaks > 1) {
        for (int k = 0; k < amountOfPeaks - 1; k++) {
            if (outPeaks) outPeaks[k] = outPeaks[k + 1];
            if (timeOfPeaks)
                timeOfPeaks[k] = (double)(timeOfPeaks[k + 1] - timeOfPeaks[k]) * h;
        }
        amountOfPeaks = amountOfPeaks - 1;
    } else {
        amountOfPeaks = 0;
    }
    return amountOfPeaks;
}

/* ------------------------------------------------------------------------ */
/* Per-point clustering                                                      */
/* ------------------------------------------------------------------------ */

/* Bifurcation/src/bifurcationCUDA.cu:575-583 */
static double distance_ref(double x1, double y1, double x2, double y2)
{
    if (x1 == x2 && y1 == y2) return 0;
    double dx = x2 - x1;
    double dy = y2 - y1;
    return (double)hypotf((float)dx, (float)dy);
}

/*
 * Bifurcation/src/bifurcationCUDA.cu:516-572, restated line by line with its
 * quirks: peak values are zeroed (clustering on intervals only), k never
 * reaches the last point, the neighbour "stack" reads slot P+NumNeibor-1
 * while pushes write slot P+k, and the result is cluster-1.
 * helpful must hold sizeOfHelpful >= 2*P-1 doubles.
 * Pinned by: hand-traced cases in tests/test_oracle_cluster.py.
 */
int orc_dbscan_ref(double* data, const double* intervals, double* helpful,
                   int amountOfPeaks, int sizeOfHelpful, double eps)
{
    if (amountOfPeaks <= 0) return 0;
    if (amountOfPeaks == 1) return 1;
    int cluster = 0;
    int NumNeibor = 0;
    for (int i = 0; i < sizeOfHelpful; ++i) helpful[i] = 0;
    for (int i = 0; i < amountOfPeaks; i++) data[i] = 0;
    for (int i = 0; i < amountOfPeaks; i++)
        if (NumNeibor >= 1) {
            i = (int)helpful[amountOfPeaks + NumNeibor - 1];
            helpful[amountOfPeaks + NumNeibor - 1] = 0;
            NumNeibor = NumNeibor - 1;
            for (int k = 0; k < amountOfPeaks - 1; k++) {
                if (i != k && helpful[k] == 0) {
                    if (distance_ref(data[i], intervals[i], data[k], intervals[k]) < eps) {
                        helpful[k] = cluster;
                        helpful[amountOfPeaks + k] = k;
                        ++NumNeibor;
                    }
                }
            }
        } else if (helpful[i] == 0) {
            NumNeibor = 0;
            ++cluster;
            helpful[i] = cluster;
            for (int k = 0; k < amountOfPeaks - 1; k++) {
                if (i != k && helpful[k] == 0) {
                    if (distance_ref(data[i], intervals[i], data[k], intervals[k]) < eps) {
                        helpful[k] = cluster;
                        helpful[amountOfPeaks + k] = k;
                        ++NumNeibor;
                    }
                }
            }
        }
    return cluster - 1;
}

static int cmp_double(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/*
 * 1-D DBSCAN with minPts = 1 (every point is core): the clusters are the
 * connected components of |u - v| <= eps, i.e. 1 + #{sorted gaps > eps}.
 * Pinned by: sklearn.cluster.DBSCAN(eps, min_samples=1) fixtures.
 */
int orc_dbscan1d_count(const double* values, int n, double eps)
{
    if (n <= 0) return 0;
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    if (!v) return -1;
    memcpy(v, values, sizeof(double) * (size_t)n);
    qsort(v, (size_t)n, sizeof(double), cmp_double);
    int c = 1;
    for (int k = 1; k < n; ++k)
        if (v[k] - v[k - 1] > eps) ++c;
    free(v);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Bifurcation drivers                                                       */
/* ------------------------------------------------------------------------ */

static int horizon(double T, double h, int sem)
{
    /* Reference: int(T/h) truncation (bifurcationCUDA.cu:60-62). */
    return sem == ORC_SEM_REFERENCE ? (int)(T / h) : (int)llround(T / h);
}

static int record_len(double tMax, double h, int preScaller, int sem)
{
    return sem == ORC_SEM_REFERENCE ? (int)(tMax / h / preScaller)
                                    : (int)llround(tMax / h / preScaller);
}

/* One bifurcation point: Bifurcation/src/bifurcationCUDA.cu:258-346.
 * Returns the flag; data[B] receives the recorded samples. */
static int bif_point(int sys, int sem, const double* ic, int nx,
                     const double* a, double h, int S, int B, int writableVar,
                     double maxValue, double* data)
{
    double X[ORC_MAX_NX];
    int chk = sem == ORC_SEM_REFERENCE ? ORC_CHK_BIF_REF : ORC_CHK_CONFIG;
    for (int i = 0; i < nx; ++i) X[i] = ic[i];
    int flag = orc_loop(sys, X, a, h, S, 1, writableVar, maxValue, chk, NULL);
    if (flag != 1) return 0;     /* bifurcationCUDA.cu:334-342 */
    return orc_loop(sys, X, a, h, B, 1, writableVar, maxValue, chk, data);
}

/* Bifurcation/src/bifurcationCUDA.cu:587-783 (values, not the CSV). */
long long orc_bif1d(int sys, int sem, double tMax, int nPts, double h,
                    const double* ic, int nx, const double* ranges,
                    const int* idx, int writableVar, double maxValue,
                    double transientTime, const double* values, int na,
                    int preScaller, int* counts, double* first,
                    double* peaks, double* intervals, long long capacity)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 1) return -2;
    int B = record_len(tMax, h, preScaller, sem);
    int S = horizon(transientTime, h, sem);
    if (B < 1) return -2;
    double* data = (double*)malloc(sizeof(double) * (size_t)B * 3);
    if (!data) return -3;
    double* pk = data + B;
    double* iv = data + 2 * (size_t)B;
    double a[ORC_MAX_NX];
    long long total = 0;
    for (int p = 0; p < nPts; ++p) {
        for (int i = 0; i < na; ++i) a[i] = values[i];
        a[idx[0]] = orc_grid_value(p % nPts, nPts, ranges[0], ranges[1]);
        int flag = bif_point(sys, sem, ic, nx, a, h, S, B, writableVar, maxValue, data);
        first[p] = 0;
        if (flag == -1) {
            counts[p] = -1;
            first[p] = data[0];
        } else if (flag == 0) {
            counts[p] = 0;
        } else {
            int P = orc_peak_finder(data, B, pk, iv, h * (double)preScaller);
            counts[p] = P;
            for (int k = 0; k < P; ++k) {
                if (total < capacity) {
                    peaks[total] = pk[k];
                    intervals[total] = iv[k];
                }
                ++total;
            }
        }
    }
    free(data);
    return total > capacity ? -1 : total;
}

/* Bifurcation/src/bifurcationCUDA.cu:43-256 (values, not the CSV).
 * REFERENCE: quirky dbscan on (zeroed peaks, intervals); CONFIG: 1-D DBSCAN
 * cluster count of the kept peak values. */
int orc_bif2d(int sys, int sem, double tMax, int nPts, double h,
              const double* ic, int nx, const double* ranges, const int* idx,
              int writableVar, double maxValue, double transientTime,
              const double* values, int na, int preScaller, double eps,
              int* out, int row_begin, int row_end)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 1) return -2;
    int B = record_len(tMax, h, preScaller, sem);
    int S = horizon(transientTime, h, sem);
    if (B < 1) return -2;
    int err = 0;
#pragma omp parallel
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)B * 4);
        if (!data) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 4)
        for (long long p = (long long)row_begin * nPts; p < (long long)row_end * nPts; ++p) {
            if (!data) continue;
            double* pk = data + B;
            double* iv = data + 2 * (size_t)B;
            double* helpful = data + 3 * (size_t)B;
            double a[ORC_MAX_NX];
            for (int i = 0; i < na; ++i) a[i] = values[i];
            a[idx[0]] = orc_grid_value((int)(p % nPts), nPts, ranges[0], ranges[1]);
            a[idx[1]] = orc_grid_value((int)(p / nPts), nPts, ranges[2], ranges[3]);
            int flag = bif_point(sys, sem, ic, nx, a, h, S, B, writableVar, maxValue, data);
            int r;
            if (flag == -1) r = -1;
            else if (flag == 0) r = 0;
            else {
                int P = orc_peak_finder(data, B, pk, iv, h * (double)preScaller);
                if (sem == ORC_SEM_REFERENCE)
                    r = orc_dbscan_ref(pk, iv, helpful, P, B, eps);
                else
                    r = orc_dbscan1d_count(pk, P, eps);
            }
            out[p - (long long)row_begin * nPts] = r;
        }
        free(data);
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* LLE (Benettin)                                                            */
/* ------------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014), the public-domain mixer. */
uint64_t orc_splitmix64(uint64_t* state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/*
 * Seeded unit vector for the Benettin perturbation (replaces curand XORWOW at
 * LLE/src/LLECUDA.cu:88-101, whose output cannot be reproduced here).
 * Stream state = seed ^ (point+1)*0xD1B54A32D192ED03; draws map to exact
 * uniforms on [-1,1) as (int64)(z>>10 - 2^53) * 2^-53; rejection to the unit
 * ball (1e-6 < r2 <= 1, at most 64 tries) then division by sqrt(r2).
 * Pinned by: |v| = 1 within 2 ulp and uniform-direction moments in tests.
 */
void orc_unit_vector(uint64_t seed, long long point, int nx, double* v)
{
    uint64_t st = seed ^ ((uint64_t)(point + 1) * 0xD1B54A32D192ED03ULL);
    for (int tr = 0; tr < 64; ++tr) {
        double r2 = 0;
        for (int i = 0; i < nx; ++i) {
            uint64_t z = orc_splitmix64(&st);
            v[i] = (double)((int64_t)(z >> 10) - ((int64_t)1 << 53)) * 0x1p-53;
        }
        for (int i = 0; i < nx; ++i) r2 = r2 + v[i] * v[i];
        if (r2 > 1e-6 && r2 <= 1.0) {
            double norm = sqrt(r2);
            for (int i = 0; i < nx; ++i) v[i] = v[i] / norm;
            return;
        }
    }
    for (int i = 0; i < nx; ++i) v[i] = i == 0 ? 1.0 : 0.0;
}

/*
 * One LLE point: LLE/src/LLECUDA.cu:56-112 (transient + perturbation) and
 * :114-166 (Benettin windows).  d = |X - Xp| / eps; sum log(d);
 * Xp = X - (X - Xp) / d; result sum / tMax.
 * REFERENCE windows: int(tMax/h)/int(NT/h) + 2 (LLECUDA.cu:227,147);
 * CONFIG windows: llround(tMax/h)/llround(NT/h).  log is the FP64 log.
 * Pinned by: linear diagonal system (LLE = ln R(h lambda)/h), stable
 * equilibrium, periodic orbit ~0 and published Lorenz/Rossler values.
 */
double orc_lle_point(int sys, int sem, double tMax, double NT, double h,
                     double eps, const double* ic, int nx, const double* a,
                     double transientTime, uint64_t seed, long long point)
{
    double X[ORC_MAX_NX], Xp[ORC_MAX_NX], v[ORC_MAX_NX];
    int S = horizon(transientTime, h, sem);
    int NTp = horizon(NT, h, sem);
    int all = horizon(tMax, h, sem);
    if (NTp <= 0) return NAN;
    int W = sem == ORC_SEM_REFERENCE ? all / NTp + 2 : all / NTp;
    for (int i = 0; i < nx; ++i) X[i] = ic[i];
    orc_steps(sys, X, a, h, S);
    orc_unit_vector(seed, point, nx, v);
    for (int i = 0; i < nx; ++i) {
        double d = eps * v[i];
        Xp[i] = X[i] + d;
    }
    double sum = 0.0;
    for (int w = 0; w < W; ++w) {
        orc_steps(sys, X, a, h, NTp);
        orc_steps(sys, Xp, a, h, NTp);
        double d2 = 0.0;
        for (int l = 0; l < nx; ++l) {
            double diff = X[l] - Xp[l];
            d2 = d2 + diff * diff;
        }
        double dist = sqrt(d2) / eps;
        sum = sum + log(dist);
        for (int l = 0; l < nx; ++l) Xp[l] = X[l] - ((X[l] - Xp[l]) / dist);
    }
    return sum / tMax;
}

/* LLE/src/LLECUDA.cu:205-428 (values, not the CSV): out[(j-row_begin)*n + i]
 * = LLE(A_i, B_j) with A = linspace(ranges[0..1]), B = linspace(ranges[2..3]);
 * the perturbation seed index is j*n + i (the reference's idx_a + idx_b*n). */
int orc_lle2d(int sys, int sem, double tMax, double NT, int nPts, double h,
              double eps, const double* ic, int nx, const double* ranges,
              const int* idx, double transientTime, const double* values,
              int na, uint64_t seed, double* out, int row_begin, int row_end)
{
    return orc_lle2d_omp(sys, sem, tMax, NT, nPts, h, eps, ic, nx, ranges, idx,
                         transientTime, values, na, seed, out, row_begin,
                         row_end, 1, 0) < 0 ? -2 : 0;
}

int orc_lle2d_omp(int sys, int sem, double tMax, double NT, int nPts, double h,
                  double eps, const double* ic, int nx, const double* ranges,
                  const int* idx, double transientTime, const double* values,
                  int na, uint64_t seed, double* out, int row_begin,
                  int row_end, int row_stride, int threads)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 1) return -2;
    if (row_stride < 1) row_stride = 1;
    int nrows = (row_end - row_begin + row_stride - 1) / row_stride;
    int used = 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
#pragma omp for schedule(dynamic, 64)
        for (long long q = 0; q < (long long)nrows * nPts; ++q) {
            int r = (int)(q / nPts), i = (int)(q % nPts);
            int j = row_begin + r * row_stride;
            double a[ORC_MAX_NX];
            for (int k = 0; k < na; ++k) a[k] = values[k];
            a[idx[0]] = orc_grid_value(i, nPts, ranges[0], ranges[1]);
            a[idx[1]] = orc_grid_value(j, nPts, ranges[2], ranges[3]);
            out[q] = orc_lle_point(sys, sem, tMax, NT, h, eps, ic, nx, a,
                                   transientTime, seed, (long long)j * nPts + i);
        }
    }
    return used;
}

/* ------------------------------------------------------------------------ */
/* Basins                                                                    */
/* ------------------------------------------------------------------------ */

/*
 * Per-IC features for Basins/src/basinsCUDA.cu:233-314 (transient without
 * checks, then B samples with preScaller sub-steps) and :818-864 (REFERENCE:
 * avg peak / avg interval with sentinels 999 / (last sample, -1) / 1000).
 * CONFIG: f0 = mean X[0], f1 = mean X[nx-1] over the B recorded samples
 * (sequential sums), f2 = population std of the n kept writableVar peaks
 * from s0 = sum p, s1 = sum fma(p,p,.): m = s0/n, var = s1/n - m*m,
 * std = sqrt(var) if var > 0 and n >= 2, else 0; divergent points get
 * (0,0,0) and flag 0.
 * Pinned by tests/test_oracle_basins.py: the sentinels, the Lorenz
 * equilibrium closed forms, the Lorenz mirror map, and an independent
 * fsum / two-pass-std / telescoped-interval evaluation of chaotic records.
 */
int orc_basins_features(int sys, int sem, double tMax, int nPts, double h,
                        const double* ic, int nx, const double* ranges,
                        const int* idx, int writableVar, double maxValue,
                        double transientTime, const double* values, int na,
                        int preScaller, double* f0, double* f1, double* f2,
                        int* flags, int point_begin, int point_end)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2) return -2;
    if (preScaller < 1) preScaller = 1;
    int B = record_len(tMax, h, preScaller, sem);
    int S = horizon(transientTime, h, sem);
    if (B < 1) return -2;
    int err = 0;
#pragma omp parallel
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)B * 3 * (sem == ORC_SEM_REFERENCE ? 1 : 2));
        if (!data) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int p = point_begin; p < point_end; ++p) {
            if (!data) continue;
            double* pk = data + B;
            double* iv = data + 2 * (size_t)B;
            double X[ORC_MAX_NX];
            const double* a = values;
            int q = p - point_begin;
            for (int i = 0; i < nx; ++i) X[i] = ic[i];
            X[idx[0]] = orc_grid_value_pow(p, nPts, ranges[0], ranges[1], 0);
            X[idx[1]] = orc_grid_value_pow(p, nPts, ranges[2], ranges[3], 1);
            orc_steps(sys, X, a, h, S);
            if (sem == ORC_SEM_REFERENCE) {
                int flag = orc_loop(sys, X, a, h, B, preScaller, writableVar,
                                    maxValue, ORC_CHK_BASINS_REF, data);
                flags[q] = flag;
                if (flag == 0) { f0[q] = 999; f1[q] = 999; }
                else if (flag == -1) { f0[q] = data[B - 1]; f1[q] = -1.0; }
                else {
                    int P = orc_peak_finder(data, B, pk, iv, h * preScaller);
                    if (P <= 0) { f0[q] = 1000; f1[q] = 1000; }
                    else {
                        double s0 = 0, s1 = 0;
                        for (int k = 0; k < P; ++k) { s0 += pk[k]; s1 += iv[k]; }
                        f0[q] = s0 / P;
                        f1[q] = s1 / P;
                    }
                }
                if (f2) f2[q] = 0;
            } else {
                /* record all state components we need: data = X[wv], plus
                 * the two mean channels accumulated alongside. */
                double* dx0 = data + 3 * (size_t)B;
                double* dxn = data + 4 * (size_t)B;
                double Y[ORC_MAX_NX];
                double xPrev[ORC_MAX_NX];
                int flag = 1;
                for (int i = 0; i < nx; ++i) Y[i] = X[i];
                for (int s = 0; s < B && flag; ++s) {
                    for (int j = 0; j < nx; ++j) xPrev[j] = Y[j];
                    data[s] = Y[writableVar];
                    dx0[s] = Y[0];
                    dxn[s] = Y[nx - 1];
                    for (int j = 0; j < preScaller; ++j) orc_step(sys, Y, a, h);
                    if (diverged(Y[writableVar], maxValue, ORC_CHK_CONFIG)) flag = 0;
                }
                if (flag) {
                    double t = 0;
                    for (int j = 0; j < nx; ++j) t += ((Y[j] - xPrev[j]) * (Y[j] - xPrev[j]));
                    flag = fixed_point(t, ORC_CHK_CONFIG) ? -1 : 1;
                }
                flags[q] = flag;
                if (flag == 0) { f0[q] = 0; f1[q] = 0; f2[q] = 0; }
                else {
                    double s0 = 0, s1 = 0;
                    for (int s = 0; s < B; ++s) { s0 = s0 + dx0[s]; s1 = s1 + dxn[s]; }
                    f0[q] = s0 / B;
                    f1[q] = s1 / B;
                    int P = orc_peak_finder(data, B, pk, iv, h * preScaller);
                    double p0 = 0, p1 = 0;
                    for (int k = 0; k < P; ++k) {
                        p0 = p0 + pk[k];
                        p1 = fma(pk[k], pk[k], p1);
                    }
                    f2[q] = 0.0;
                    if (P >= 2) {
                        double m = p0 / P;
                        double var = p1 / P - m * m;
                        f2[q] = var > 0 ? sqrt(var) : 0.0;
                    }
                }
            }
        }
        free(data);
    }
    return err;
}

/* Union-find helpers (path halving, union by smaller index as root). */
static int uf_find(int* par, int x)
{
    while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
    return x;
}
static void uf_union(int* par, int a, int b)
{
    a = uf_find(par, a); b = uf_find(par, b);
    if (a == b) return;
    if (a < b) par[b] = a; else par[a] = b;
}

/*
 * Basins/src/basinsCUDA.cu:35-179,866-923 without its races: components of
 * the graph sqrt(dx*dx + dy*dy) <= eps over non-divergent points (flag != 0;
 * the reference never expands through a divergent point, :913-916).  A
 * component holding a fixed point (flag -1) gets a negative label, the rest
 * positive (fixed-point seeds are searched first, :86-126).  Labels are
 * numbered by smallest member index within each sign (the reference's seed
 * choice is racy, so its numbering is pinned only up to permutation).
 * O(n^2): for n up to ~2e4.  Pinned by: hand-built point sets in tests.
 */
int orc_basins_cluster_ref(const double* f0, const double* f1,
                           const int* flags, int n, double eps, int* labels)
{
    int* par = (int*)malloc(sizeof(int) * (size_t)n * 3);
    if (!par) return -3;
    int* neg = par + n;
    int* num = par + 2 * (size_t)n;
    for (int i = 0; i < n; ++i) { par[i] = i; neg[i] = 0; num[i] = 0; }
    for (int i = 0; i < n; ++i) {
        if (flags[i] == 0) continue;
        for (int j = i + 1; j < n; ++j) {
            if (flags[j] == 0) continue;
            double dx = f0[i] - f0[j];
            double dy = f1[i] - f1[j];
            if (sqrt(dx * dx + dy * dy) <= eps) uf_union(par, i, j);
        }
    }
    for (int i = 0; i < n; ++i)
        if (flags[i] == -1) neg[uf_find(par, i)] = 1;
    int npos = 0, nneg = 0;
    for (int i = 0; i < n; ++i) {
        if (flags[i] == 0) { labels[i] = 0; continue; }
        int r = uf_find(par, i);
        if (neg[r]) {
            /* number negatives by smallest fixed-point index */
            if (flags[i] == -1 && num[r] == 0) num[r] = -(++nneg);
        } else if (num[r] == 0) {
            num[r] = ++npos;
        }
    }
    for (int i = 0; i < n; ++i)
        labels[i] = flags[i] == 0 ? 0 : num[uf_find(par, i)];
    free(par);
    return 0;
}

/*
 * Deterministic fixed-order sum shared with the HIP reduction: the entries
 * (masked-out ones and the padding contribute +0) padded to P = max(1024,
 * the power of two >= n) and summed by the complete pairwise tree over
 * adjacent pairs: level 1 adds (v[2t], v[2t+1]), level 2 the sums of
 * (v[4t..4t+1], v[4t+2..4t+3]), and so on up to the whole range.  Every
 * aligned run of 2^k entries is a subtree, so the device may form the sums
 * of 256- or 1024-entry runs in its blocks and combine them by the same tree.
 * Pinned by: math.fsum accuracy and an independent recursive evaluation of
 * the tree (tests/test_oracle_cluster.py).
 */
double orc_det_sum(const double* x, const int* mask, long long n)
{
    long long P = 1024;
    while (P < n) P <<= 1;
    /* level 1 on the fly (1024-entry runs), then the upper levels */
    const long long nr = P / 1024;
    double* run = (double*)malloc(sizeof(double) * (size_t)nr);
    if (!run) return NAN;
    double s[1024];
    for (long long r = 0; r < nr; ++r) {
        for (int t = 0; t < 1024; ++t) {
            const long long i = r * 1024 + t;
            s[t] = (i < n && (!mask || mask[i] != 0)) ? x[i] : 0.0;
        }
        for (int w = 1024; w > 1; w >>= 1)
            for (int t = 0; t < w / 2; ++t) s[t] = s[2 * t] + s[2 * t + 1];
        run[r] = s[0];
    }
    for (long long w = nr; w > 1; w >>= 1)
        for (long long t = 0; t < w / 2; ++t) run[t] = run[2 * t] + run[2 * t + 1];
    const double tot = run[0];
    free(run);
    return tot;
}

/* Z-score the three CONFIG features over points with flag != 0, in place:
 * mean = S/N, std = sqrt(sum((x-mean)^2)/N) (two passes, det_sum order);
 * z = (x - mean) / std, or 0 when std == 0.  Excluded points are left. */
int orc_zscore(double* f0, double* f1, double* f2, const int* flags, long long n,
               double* moments)
{
    double* f[3] = {f0, f1, f2};
    long long N = 0;
    for (long long i = 0; i < n; ++i) N += flags[i] != 0;
    if (N == 0) return 0;
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
    if (!tmp) return -3;
    for (int k = 0; k < 3; ++k) {
        double mean = orc_det_sum(f[k], flags, n) / (double)N;
        for (long long i = 0; i < n; ++i) {
            double d = f[k][i] - mean;
            tmp[i] = d * d;
        }
        double sd = sqrt(orc_det_sum(tmp, flags, n) / (double)N);
        if (moments) { moments[2 * k] = mean; moments[2 * k + 1] = sd; }
        for (long long i = 0; i < n; ++i) {
            if (flags[i] == 0) continue;
            f[k][i] = sd > 0 ? (f[k][i] - mean) / sd : 0.0;
        }
    }
    free(tmp);
    return 0;
}

/*
 * Brute-force DBSCAN (O(n^2)) with the build's deterministic rules: a point is
 * core when #{q : sqrt(sum_k (p_k - q_k)^2) <= eps} >= minPts (self counted,
 * squares summed in coordinate order); core points within eps are joined; a
 * non-core point within eps of a core point takes the cluster of the
 * smallest-index such core point; noise and excluded points get 0; clusters
 * are numbered 1..K by their smallest core index.  pts is [n][dim].
 * Pinned by: sklearn.cluster.DBSCAN core-sample partition equality.
 */
int orc_dbscan_bruteforce(const double* pts, int n, int dim, double eps,
                          int minPts, const int* include, int* labels)
{
    int* par = (int*)malloc(sizeof(int) * (size_t)n * 3);
    if (!par) return -3;
    int* core = par + n;
    int* num = par + 2 * (size_t)n;
#define ORC_D(i, j, out)                                               \
    do {                                                               \
        double d2_ = 0.0;                                              \
        for (int k_ = 0; k_ < dim; ++k_) {                             \
            double df_ = pts[(size_t)(i) * dim + k_] - pts[(size_t)(j) * dim + k_]; \
            d2_ = d2_ + df_ * df_;                                     \
        }                                                              \
        (out) = sqrt(d2_) <= eps;                                      \
    } while (0)
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        core[i] = 0;
        if (include && !include[i]) continue;
        int c = 0;
        for (int j = 0; j < n && c < minPts; ++j) {
            if (include && !include[j]) continue;
            int w;
            ORC_D(i, j, w);
            c += w;
        }
        core[i] = c >= minPts;
    }
    for (int i = 0; i < n; ++i) { par[i] = i; num[i] = 0; }
    for (int i = 0; i < n; ++i) {
        if (!core[i]) continue;
        for (int j = i + 1; j < n; ++j) {
            if (!core[j]) continue;
            int w;
            ORC_D(i, j, w);
            if (w) uf_union(par, i, j);
        }
    }
    int K = 0;
    for (int i = 0; i < n; ++i)
        if (core[i]) {
            int r = uf_find(par, i);
            if (num[r] == 0) num[r] = ++K;
        }
    /* core labels first (sequential: uf_find compresses paths) */
    for (int i = 0; i < n; ++i) labels[i] = core[i] ? num[uf_find(par, i)] : 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        if (core[i] || (include && !include[i])) continue;
        for (int j = 0; j < n; ++j) {
            if (!core[j]) continue;
            int w;
            ORC_D(i, j, w);
            if (w) { labels[i] = labels[j]; break; }
        }
    }
#undef ORC_D
    free(par);
    return K;
}

/*
 * The same clustering as orc_dbscan_bruteforce in O(n) neighbour work, for
 * full-size label checks (millions of points): points binned into cells of
 * side s = eps/sqrt(dim) * (1 - 1e-7) (two points of one cell lie within
 * eps), a point's eps-neighbours all lie in the 5^dim cells around its own,
 * and a neighbour cell is taken whole or skipped when its members' exact box
 * lies wholly within / beyond eps (the gap and far distances are formed like
 * the point distance, so the verdicts agree with it point by point).  Same
 * rules and numbering as the brute force.  Returns K, -2 when a coordinate
 * range is too wide for the cell keys, -3 on allocation failure.
 * Pinned by: equality with orc_dbscan_bruteforce (tests/test_oracle_cluster.py).
 */
typedef struct { unsigned long long key; int idx; } orc_ki;

static int orc_cmp_ki(const void* a, const void* b)
{
    const orc_ki* x = (const orc_ki*)a;
    const orc_ki* y = (const orc_ki*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

typedef struct {
    const double* pts;
    int dim;
    double eps;
    int ncell;
    const unsigned long long* ckey;
    const int* cstart;   /* [ncell + 1] into ord */
    const orc_ki* ord;
    const double* cmin;  /* [ncell][3] */
    const double* cmax;
} orc_grid;

static int orc_within(const orc_grid* g, int p, int q)
{
    double d2 = 0.0;
    for (int k = 0; k < g->dim; ++k) {
        double df = g->pts[(size_t)p * g->dim + k] - g->pts[(size_t)q * g->dim + k];
        d2 = d2 + df * df;
    }
    return sqrt(d2) <= g->eps;
}

/* 0: no member within eps of p; 2: every member within eps; 1: undecided. */
static int orc_box_relation(const orc_grid* g, int p, int c)
{
    double g2 = 0.0, f2 = 0.0;
    for (int k = 0; k < g->dim; ++k) {
        double lo = g->cmin[3 * c + k], hi = g->cmax[3 * c + k];
        double v = g->pts[(size_t)p * g->dim + k];
        double gap = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        double far = fmax(v - lo, hi - v);
        g2 = g2 + gap * gap;
        f2 = f2 + far * far;
    }
    if (sqrt(g2) > g->eps) return 0;
    if (sqrt(f2) <= g->eps) return 2;
    return 1;
}

static int orc_find_cell(const orc_grid* g, unsigned long long key)
{
    int a = 0, b = g->ncell;
    while (a < b) {
        int m = (a + b) / 2;
        if (g->ckey[m] < key) a = m + 1;
        else b = m;
    }
    return a < g->ncell && g->ckey[a] == key ? a : -1;
}

static unsigned long long orc_pack(int dim, const long long* c)
{
    if (dim == 3) return (unsigned long long)c[0] | ((unsigned long long)c[1] << 21) | ((unsigned long long)c[2] << 42);
    return (unsigned long long)c[0] | ((unsigned long long)c[1] << 31);
}

/* Cells of the 5^dim block around cell `c0` coordinates: out[] indices (-1
 * for none); returns the count written. */
static int orc_neighbours(const orc_grid* g, const long long* c0, int* out)
{
    int m = 0;
    int rz = g->dim == 3 ? 2 : 0;
    for (int dz = -rz; dz <= rz; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                long long c[3] = {c0[0] + dx, c0[1] + dy, c0[2] + dz};
                if (c[0] < 0 || c[1] < 0 || c[2] < 0) continue;
                int k = orc_find_cell(g, orc_pack(g->dim, c));
                if (k >= 0) out[m++] = k;
            }
    return m;
}

int orc_dbscan_grid(const double* pts, int n, int dim, double eps, int minPts,
                    const int* include, int* labels)
{
    if (dim != 2 && dim != 3) return -1;
    double s = eps / sqrt((double)dim) * (1.0 - 1e-7);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int ninc = 0;
    for (int i = 0; i < n; ++i) {
        labels[i] = 0;
        if (include && !include[i]) continue;
        ++ninc;
        for (int k = 0; k < dim; ++k) {
            double v = pts[(size_t)i * dim + k];
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
    }
    if (ninc == 0) return 0;
    double lim = dim == 3 ? (double)(1 << 21) - 4 : 2147483000.0;
    for (int k = 0; k < dim; ++k)
        if (floor((hi[k] - lo[k]) / s) + 1 >= lim) return -2;
    orc_ki* ord = (orc_ki*)malloc(sizeof(orc_ki) * (size_t)ninc);
    int* cell_of = (int*)malloc(sizeof(int) * (size_t)n);
    int* core = (int*)calloc((size_t)n, sizeof(int));
    int* par = (int*)malloc(sizeof(int) * (size_t)n);
    int* num = (int*)calloc((size_t)n, sizeof(int));
    if (!ord || !cell_of || !core || !par || !num) {
        free(ord); free(cell_of); free(core); free(par); free(num);
        return -3;
    }
    int m = 0;
    for (int i = 0; i < n; ++i) {
        cell_of[i] = -1;
        if (include && !include[i]) continue;
        long long c[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c[k] = (long long)floor((pts[(size_t)i * dim + k] - lo[k]) / s);
        ord[m].key = orc_pack(dim, c);
        ord[m].idx = i;
        ++m;
    }
    qsort(ord, (size_t)ninc, sizeof(orc_ki), orc_cmp_ki);
    int ncell = 0;
    for (int t = 0; t < ninc; ++t)
        if (t == 0 || ord[t].key != ord[t - 1].key) ++ncell;
    unsigned long long* ckey = (unsigned long long*)malloc(sizeof(unsigned long long) * (size_t)ncell);
    int* cstart = (int*)malloc(sizeof(int) * (size_t)(ncell + 1));
    double* cmin = (double*)malloc(sizeof(double) * 3 * (size_t)ncell);
    double* cmax = (double*)malloc(sizeof(double) * 3 * (size_t)ncell);
    int* cmincore = (int*)malloc(sizeof(int) * (size_t)ncell);
    if (!ckey || !cstart || !cmin || !cmax || !cmincore) {
        free(ord); free(cell_of); free(core); free(par); free(num);
        free(ckey); free(cstart); free(cmin); free(cmax); free(cmincore);
        return -3;
    }
    int c = -1;
    for (int t = 0; t < ninc; ++t) {
        if (t == 0 || ord[t].key != ord[t - 1].key) {
            ++c;
            ckey[c] = ord[t].key;
            cstart[c] = t;
            for (int k = 0; k < 3; ++k) { cmin[3 * c + k] = INFINITY; cmax[3 * c + k] = -INFINITY; }
        }
        int p = ord[t].idx;
        cell_of[p] = c;
        for (int k = 0; k < dim; ++k) {
            double v = pts[(size_t)p * dim + k];
            cmin[3 * c + k] = fmin(cmin[3 * c + k], v);
            cmax[3 * c + k] = fmax(cmax[3 * c + k], v);
        }
    }
    cstart[ncell] = ninc;
    orc_grid g = {pts, dim, eps, ncell, ckey, cstart, ord, cmin, cmax};

    /* core points: whole dense cells, counted neighbourhoods elsewhere */
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < ninc; ++t) {
        int p = ord[t].idx, cp = cell_of[p];
        if (cstart[cp + 1] - cstart[cp] >= minPts) { core[p] = 1; continue; }
        long long c0[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c0[k] = (long long)floor((pts[(size_t)p * dim + k] - lo[k]) / s);
        int nb[125];
        int nn = orc_neighbours(&g, c0, nb);
        long long cnt = 0;
        for (int j = 0; j < nn && cnt < minPts; ++j) {
            int rel = orc_box_relation(&g, p, nb[j]);
            if (rel == 2) cnt += cstart[nb[j] + 1] - cstart[nb[j]];
            else if (rel == 1)
                for (int u = cstart[nb[j]]; u < cstart[nb[j] + 1] && cnt < minPts; ++u)
                    cnt += orc_within(&g, p, ord[u].idx);
        }
        core[p] = cnt >= minPts;
    }
    /* smallest core index per cell (runs are sorted by index) */
    for (int k = 0; k < ncell; ++k) {
        cmincore[k] = -1;
        for (int u = cstart[k]; u < cstart[k + 1]; ++u)
            if (core[ord[u].idx]) { cmincore[k] = ord[u].idx; break; }
    }
    for (int i = 0; i < n; ++i) par[i] = i;
    /* cores of one cell are mutually within eps */
    for (int k = 0; k < ncell; ++k) {
        if (cmincore[k] < 0) continue;
        for (int u = cstart[k]; u < cstart[k + 1]; ++u)
            if (core[ord[u].idx]) uf_union(par, ord[u].idx, cmincore[k]);
    }
    /* one witness pair per pair of neighbouring core cells */
    for (int a = 0; a < ncell; ++a) {
        if (cmincore[a] < 0) continue;
        int pa = ord[cstart[a]].idx;
        long long c0[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c0[k] = (long long)floor((pts[(size_t)pa * dim + k] - lo[k]) / s);
        int nb[125];
        int nn = orc_neighbours(&g, c0, nb);
        for (int j = 0; j < nn; ++j) {
            int b = nb[j];
            if (b <= a || cmincore[b] < 0) continue;
            if (uf_find(par, cmincore[a]) == uf_find(par, cmincore[b])) continue;
            int done = 0;
            for (int u = cstart[a]; u < cstart[a + 1] && !done; ++u) {
                int p = ord[u].idx;
                if (!core[p]) continue;
                int rel = orc_box_relation(&g, p, b);
                if (rel == 0) continue;
                if (rel == 2) { uf_union(par, p, cmincore[b]); done = 1; break; }
                for (int w = cstart[b]; w < cstart[b + 1]; ++w) {
                    int q = ord[w].idx;
                    if (core[q] && orc_within(&g, p, q)) { uf_union(par, p, q); done = 1; break; }
                }
            }
        }
    }
    /* clusters numbered by smallest core index */
    int K = 0;
    for (int i = 0; i < n; ++i)
        if (core[i]) {
            int r = uf_find(par, i);
            if (num[r] == 0) num[r] = ++K;
        }
    for (int i = 0; i < n; ++i)
        if (core[i]) labels[i] = num[uf_find(par, i)];
    /* border points: the cluster of the smallest-index core neighbour */
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < ninc; ++t) {
        int p = ord[t].idx;
        if (core[p]) continue;
        long long c0[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) c0[k] = (long long)floor((pts[(size_t)p * dim + k] - lo[k]) / s);
        int nb[125];
        int nn = orc_neighbours(&g, c0, nb);
        int best = n;
        for (int j = 0; j < nn; ++j) {
            int b = nb[j];
            if (cmincore[b] < 0) continue;
            int rel = orc_box_relation(&g, p, b);
            if (rel == 0) continue;
            if (rel == 2) { if (cmincore[b] < best) best = cmincore[b]; continue; }
            for (int w = cstart[b]; w < cstart[b + 1]; ++w) {
                int q = ord[w].idx;
                if (q >= best) break;
                if (core[q] && orc_within(&g, p, q)) { best = q; break; }
            }
        }
        labels[p] = best < n ? labels[best] : 0;
    }
    free(ord); free(cell_of); free(core); free(par); free(num);
    free(ckey); free(cstart); free(cmin); free(cmax); free(cmincore);
    return K;
}

/* ------------------------------------------------------------------------ */
/* Legacy entry points (debri/, namespace old_library)                      */
/* ------------------------------------------------------------------------ */

/* Horizon of the legacy drivers: `int x = T / h` truncation
 * (debri/hostLibrary.cu:3128-3136) or llround in CONFIG semantics. */
static long long legacy_count(double q, int sem)
{
    return sem == ORC_SEM_REFERENCE ? (long long)q : llround(q);
}

/* debri/cudaLibrary.cu:472-524 (loopCalculateDiscreteModel) as the LLE and
 * LS kernels call it (preScaller 1, writableVar 0): n steps, stopping early
 * after a step that leaves X[0] NaN/inf or fabsf(X[0]) > maxValue
 * (maxValue != 0).  Its return value is discarded by every caller.
 * Pinned by: divergence cases in tests/test_oracle_legacy.py. */
static void legacy_run(int sys, double* X, const double* a, double h, long long n, double maxValue)
{
    for (long long i = 0; i < n; ++i) {
        orc_step(sys, X, a, h);
        if (diverged(X[0], maxValue, ORC_CHK_BASINS_REF)) return;
    }
}

/* Swept values of legacy point idx (debri/cudaLibrary.cu:1182-1186,
 * getValueByIdx with valueNumber = dimension index). */
static void legacy_point(int idx, int nPts, int dims, const double* ranges, const int* mut,
                         double* target)
{
    for (int d = 0; d < dims; ++d)
        target[mut[d]] = orc_grid_value_pow(idx, nPts, ranges[2 * d], ranges[2 * d + 1], d);
}

/*
 * debri/cudaLibrary.cu:1897-2031 (LLEKernelCUDA, form 0: parameters swept)
 * and :2034-2137 (LLEKernelICCUDA, form 1: initial conditions swept).
 * Windows W = int(tMax/NT), NT/h steps each, on X then on the perturbed Y;
 * form 0: d = sqrt(sum((1/eps)(x-y))^2), Y = X - (X-Y)/d;
 * form 1: d = sqrt(sum (x-y)^2)/eps, Y = X - (X-Y)*(d != 0 ? 1/d : d);
 * result sum(log d)/tMax.  The initial direction is the seeded unit vector
 * (orc_unit_vector at the point index) instead of curand / sinf.
 * Pinned by: the linear diagonal system closed form, a stable equilibrium,
 * form 0 == form 1 where both are defined, and the LLE module on the same
 * trajectory (tests/test_oracle_legacy.py).
 */
double orc_legacy_lle_point(int sys, int sem, int form, double tMax, double NT, double h,
                            double eps, const double* X0, int nx, const double* a,
                            double maxValue, double transientTime, uint64_t seed,
                            long long point)
{
    double X[ORC_MAX_NX], Y[ORC_MAX_NX], v[ORC_MAX_NX];
    long long S = legacy_count(transientTime / h, sem);
    long long NTp = legacy_count(NT / h, sem);
    long long W = legacy_count(tMax / NT, sem);
    for (int i = 0; i < nx; ++i) X[i] = X0[i];
    legacy_run(sys, X, a, h, S, maxValue);
    orc_unit_vector(seed, point, nx, v);
    for (int i = 0; i < nx; ++i) Y[i] = X[i] + v[i] * eps;
    double result = 0;
    for (long long w = 0; w < W; ++w) {
        legacy_run(sys, X, a, h, NTp, maxValue);
        legacy_run(sys, Y, a, h, NTp, maxValue);
        double t = 0;
        if (form == 0) {
            for (int l = 0; l < nx; ++l) {
                double t2 = 1 / eps;
                t2 = t2 * (X[l] - Y[l]);
                t += t2 * t2;
            }
            t = sqrt(t);
            result += log(t);
            for (int j = 0; j < nx; ++j) Y[j] = X[j] - ((X[j] - Y[j]) / t);
        } else {
            for (int l = 0; l < nx; ++l) t += (X[l] - Y[l]) * (X[l] - Y[l]);
            t = sqrt(t) / eps;
            result += log(t);
            if (t != 0) t = 1 / t;
            for (int j = 0; j < nx; ++j) Y[j] = X[j] - ((X[j] - Y[j]) * t);
        }
    }
    return result / tMax;
}

/* Legacy LLE over a 1-D or 2-D sweep of parameters (form 0) or initial
 * conditions (form 1): out[idx], idx = k0 + nPts*k1 (debri/hostLibrary.cu:
 * 3111-3325 LLE1D, :3328-3543 LLE1DIC, :3692-3827 LLE2DIC). */
int orc_legacy_lle(int sys, int sem, int form, int dims, double tMax, double NT, int nPts,
                   double h, double eps, const double* ic, int nx, const double* ranges,
                   const int* idx, double maxValue, double transientTime,
                   const double* values, int na, uint64_t seed, double* out,
                   int point_begin, int point_end)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2) return -2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = point_begin; p < point_end; ++p) {
        double X[ORC_MAX_NX], a[ORC_MAX_NX];
        for (int i = 0; i < nx; ++i) X[i] = ic[i];
        for (int i = 0; i < na; ++i) a[i] = values[i];
        legacy_point(p, nPts, dims, ranges, idx, form == 0 ? a : X);
        out[p - point_begin] = orc_legacy_lle_point(sys, sem, form, tMax, NT, h, eps, X, nx, a,
                                                    maxValue, transientTime, seed, p);
    }
    return 0;
}

/* debri/cudaLibrary.cu:2141-2187 (projectionOperator, gramSchmidtProcess):
 * classical Gram-Schmidt of the n rows of a into b (each row minus its
 * projections, numerator from the original row, onto the earlier
 * unnormalised rows), then every row normalised; den[i] = row norm.
 * Pinned by: orthonormality, span and the first row's direction
 * (tests/test_oracle_legacy.py). */
void orc_gram_schmidt(const double* a, double* b, int n, double* den)
{
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) b[j + i * n] = a[j + i * n];
        for (int j = 0; j < i; ++j) {
            double num = 0, dn = 0;
            for (int k = 0; k < n; ++k) {
                num += a[i * n + k] * b[j * n + k];
                dn += b[j * n + k] * b[j * n + k];
            }
            double frac = dn == 0 ? 0 : num / dn;
            for (int k = 0; k < n; ++k) b[i * n + k] -= frac * b[j * n + k];
        }
    }
    for (int i = 0; i < n; ++i) {
        double d = 0;
        for (int j = 0; j < n; ++j) d += b[i * n + j] * b[i * n + j];
        d = sqrt(d);
        for (int j = 0; j < n; ++j) b[i * n + j] = d == 0 ? 0 : b[i * n + j] / d;
        if (den) den[i] = d;
    }
}

/*
 * debri/cudaLibrary.cu:2189-2331 (LSKernelCUDA): Lyapunov spectrum with nx
 * perturbed trajectories.  Initial directions: nx seeded unit vectors
 * (orc_unit_vector at point*4 + j, instead of the sinf rows), Gram-Schmidt,
 * Y_j = Y_j*eps + X.  Per window: X and every Y_j advance NT/h steps (early
 * stop as legacy_run), D_k = Y_k - X, Gram-Schmidt(D -> Z, den),
 * res[k] += log(den[k]/eps), Y_k = X + Z_k*eps.  out[k] = res[k]/tMax.
 * Pinned by: the linear diagonal system (spectrum = ln R(h lambda_k)/h in
 * order), sum of exponents = trace for Lorenz (-(sigma+1+beta)), largest
 * exponent vs the LLE oracle (tests/test_oracle_legacy.py).
 */
int orc_legacy_ls_point(int sys, int sem, double tMax, double NT, double h, double eps,
                        const double* X0, int nx, const double* a, double maxValue,
                        double transientTime, uint64_t seed, long long point, double* out)
{
    double X[ORC_MAX_NX], Y[16] = {0}, Z[16] = {0}, den[4] = {0}, res[4] = {0};
    if (nx > 4) return -2;
    long long S = legacy_count(transientTime / h, sem);
    long long NTp = legacy_count(NT / h, sem);
    long long W = legacy_count(tMax / NT, sem);
    for (int i = 0; i < nx; ++i) {
        X[i] = X0[i];
        res[i] = 0;
    }
    for (int j = 0; j < nx; ++j) orc_unit_vector(seed, point * 4 + j, nx, Z + j * nx);
    legacy_run(sys, X, a, h, S, maxValue);
    orc_gram_schmidt(Z, Y, nx, NULL);
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < nx; ++i) Y[j * nx + i] = Y[j * nx + i] * eps + X[i];
    for (long long w = 0; w < W; ++w) {
        legacy_run(sys, X, a, h, NTp, maxValue);
        for (int j = 0; j < nx; ++j) legacy_run(sys, Y + j * nx, a, h, NTp, maxValue);
        for (int k = 0; k < nx; ++k)
            for (int l = 0; l < nx; ++l) Y[k * nx + l] = Y[k * nx + l] - X[l];
        orc_gram_schmidt(Y, Z, nx, den);
        for (int k = 0; k < nx; ++k) {
            res[k] += log(den[k] / eps);
            for (int j = 0; j < nx; ++j) Y[k * nx + j] = X[j] + Z[k * nx + j] * eps;
        }
    }
    for (int i = 0; i < nx; ++i) out[i] = res[i] / tMax;
    return 0;
}

/* Legacy LS over a 1-D or 2-D parameter sweep: out[idx*nx + k]
 * (debri/hostLibrary.cu:3829-3959 LS1D, :3961-4113 LS2D). */
int orc_legacy_ls(int sys, int sem, int dims, double tMax, double NT, int nPts, double h,
                  double eps, const double* ic, int nx, const double* ranges, const int* idx,
                  double maxValue, double transientTime, const double* values, int na,
                  uint64_t seed, double* out, int point_begin, int point_end)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2 || nx > 4) return -2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = point_begin; p < point_end; ++p) {
        double a[ORC_MAX_NX];
        for (int i = 0; i < na; ++i) a[i] = values[i];
        legacy_point(p, nPts, dims, ranges, idx, a);
        orc_legacy_ls_point(sys, sem, tMax, NT, h, eps, ic, nx, a, maxValue, transientTime, seed, p,
                            out + (long long)(p - point_begin) * nx);
    }
    return 0;
}

/*
 * debri/hostLibrary.cu:5094-5331 (TimeDomainCalculation) with
 * debri/cudaLibrary.cu:879-953 (calculateDiscreteModelCUDA) and :604-680
 * (loopCalculateDiscreteModel_int): per parameter point, the transient runs
 * int(transientTime/h) iterations of preScaller steps with the checks on
 * X[writableVar]; only a flag of 1 records B = int(tMax/h/preScaller)
 * samples (X[writableVar] before each iteration).  Samples after a
 * divergence, and all samples of an unrecorded point, are 0 here (the
 * reference leaves that memory uninitialised).  out[p*B + t]; flags[p].
 * Pinned by: orc_loop's pins and a hand-integrated short run.
 */
long long orc_time_domain(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                          const double* ranges, const int* idx, int writableVar, double maxValue,
                          double transientTime, const double* values, int na, int preScaller,
                          double* out, int* flags)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2 || preScaller < 1) return -2;
    long long B = legacy_count(tMax / h / preScaller, sem);
    long long S = legacy_count(transientTime / h, sem);
    for (long long i = 0; i < (long long)nPts * B; ++i) out[i] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < nPts; ++p) {
        double X[ORC_MAX_NX], a[ORC_MAX_NX];
        for (int i = 0; i < nx; ++i) X[i] = ic[i];
        for (int i = 0; i < na; ++i) a[i] = values[i];
        legacy_point(p, nPts, 1, ranges, idx, a);
        int flag = orc_loop(sys, X, a, h, (int)S, preScaller, writableVar, maxValue, ORC_CHK_BASINS_REF, NULL);
        if (flag == 1)
            flag = orc_loop(sys, X, a, h, (int)B, preScaller, writableVar, maxValue, ORC_CHK_BASINS_REF,
                            out + (long long)p * B);
        if (flags) flags[p] = flag;
    }
    return B;
}

/* One legacy IC-sweep bifurcation point: debri/cudaLibrary.cu:1022-1062
 * (calculateDiscreteModelICCUDA) with the legacy loop (:604-680).  The
 * transient, int(transientTime/h) single steps, tests X[0] for NaN/inf only
 * (maxValue 0); a flag of 1 or -1 records B samples of X[writableVar] with
 * preScaller steps each and the full checks.  Returns the flag. */
static int legacy_ic_point(int sys, int sem, double* X, const double* a, double h, int S, int B,
                           int preScaller, int writableVar, double maxValue, double* data)
{
    int chk = sem == ORC_SEM_REFERENCE ? ORC_CHK_BASINS_REF : ORC_CHK_CONFIG;
    int flag = orc_loop(sys, X, a, h, S, 1, 0, 0.0, chk, NULL);
    if (flag != 1 && flag != -1) return 0;
    return orc_loop(sys, X, a, h, B, preScaller, writableVar, maxValue, chk, data);
}

/* debri/hostLibrary.cu:1243-1492 (bifurcation1DIC, values not the CSV):
 * counts[p] = kept peaks (peakFinder, first dropped) of flag-1 points, else
 * the flag (0 or -1, which the legacy writer skips); kept peak values
 * concatenated.  Returns the total or -1 if capacity is too small.
 * Pinned by: the IC sweep equals the parameter-sweep driver on a system
 * where a parameter enters only through the state (tests). */
long long orc_legacy_bif1d_ic(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                              const double* ranges, const int* idx, int writableVar, double maxValue,
                              double transientTime, const double* values, int na, int preScaller,
                              int* counts, double* peaks, long long capacity)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2 || preScaller < 1) return -2;
    int B = record_len(tMax, h, preScaller, sem);
    int S = horizon(transientTime, h, sem);
    if (B < 1) return -2;
    double* data = (double*)malloc(sizeof(double) * (size_t)B * 3);
    if (!data) return -3;
    long long total = 0;
    for (int p = 0; p < nPts; ++p) {
        double X[ORC_MAX_NX];
        for (int i = 0; i < nx; ++i) X[i] = ic[i];
        X[idx[0]] = orc_grid_value_pow(p, nPts, ranges[0], ranges[1], 0);
        int flag = legacy_ic_point(sys, sem, X, values, h, S, B, preScaller, writableVar, maxValue, data);
        if (flag != 1) {
            counts[p] = flag;
            continue;
        }
        int P = orc_peak_finder(data, B, data + B, data + 2 * (size_t)B, h * (double)preScaller);
        counts[p] = P;
        for (int k = 0; k < P; ++k, ++total)
            if (total < capacity) peaks[total] = data[B + k];
    }
    free(data);
    return total > capacity ? -1 : total;
}

/* debri/hostLibrary.cu:2819-3109 (bifurcation2DIC, values not the CSV):
 * out[p] = -1 / 0 for flags -1 / 0, else the quirky per-point dbscan
 * (debri/cudaLibrary.cu:1593-1666, the same text as
 * Bifurcation/src/bifurcationCUDA.cu:516-572) of the kept peaks and
 * intervals (REFERENCE) or the 1-D DBSCAN count of the kept peaks (CONFIG). */
int orc_legacy_bif2d_ic(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                        const double* ranges, const int* idx, int writableVar, double maxValue,
                        double transientTime, const double* values, int na, int preScaller, double eps,
                        int* out)
{
    int snx, sna;
    if (orc_dims(sys, &snx, &sna) || nx != snx || na != sna || nPts < 2 || preScaller < 1) return -2;
    int B = record_len(tMax, h, preScaller, sem);
    int S = horizon(transientTime, h, sem);
    if (B < 1) return -2;
    int err = 0;
#pragma omp parallel
    {
        double* data = (double*)malloc(sizeof(double) * (size_t)B * 4);
        if (!data) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 4)
        for (long long p = 0; p < (long long)nPts * nPts; ++p) {
            if (!data) continue;
            double X[ORC_MAX_NX];
            for (int i = 0; i < nx; ++i) X[i] = ic[i];
            X[idx[0]] = orc_grid_value_pow((int)p, nPts, ranges[0], ranges[1], 0);
            X[idx[1]] = orc_grid_value_pow((int)p, nPts, ranges[2], ranges[3], 1);
            int flag = legacy_ic_point(sys, sem, X, values, h, S, B, preScaller, writableVar, maxValue, data);
            int r = flag;
            if (flag == 1) {
                double* pk = data + B;
                double* iv = data + 2 * (size_t)B;
                int P = orc_peak_finder(data, B, pk, iv, h * (double)preScaller);
                r = sem == ORC_SEM_REFERENCE ? orc_dbscan_ref(pk, iv, data + 3 * (size_t)B, P, B, eps)
                                             : orc_dbscan1d_count(pk, P, eps);
            }
            out[p] = r;
        }
        free(data);
    }
    return err;
}
