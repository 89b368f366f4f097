// green_dev.hpp -- device code for the free-surface Green functions and the
// Newman panel integrals, gfx950.
//
// Numerics follow the reference routines cited at each function (FP64, the
// reference's own polynomial fits and series, its REAL*4 literals kept as
// (double)(float) values).  Structure is GPU-first: no recursion (the
// adaptive Gauss-Kronrod of AQUAD runs on a bounded explicit stack), the
// epsilon-table rows live in caller-owned per-lane arrays, and every input of
// a discrete decision (FLAG, the R/h region, series lengths) is formed with
// FMA contraction off so that it takes the same branch as the FP64 reference.
//
// Acknowledgements and licences of the methods restated here:
//  * The special-function routines (error function, exponential integrals
//    E1/En/Ei, modified Bessel I0/I1/K0/K1 and Bessel J0/J1/Y0/Y1 series of
//    IK01A and JY01B) follow "Computation of Special Functions" by Shanjie
//    Zhang and Jianming Jin (Wiley, 1996).  The routines are copyrighted by
//    Shanjie Zhang and Jianming Jin, who give permission to incorporate them
//    into a user program provided that the copyright is acknowledged; it is
//    acknowledged here (as in the reference, src/FinGrnExtSubs.f90:27-31,
//    :98-102 and the headers of its other routines).
//  * The finite-depth Green function (the four R/h regions, their series
//    lengths, Wynn's epsilon extrapolation, the Gauss-Kronrod quadrature and
//    the Chebyshev image sums) follows FinGreen3D by Yingyi Liu (2013-2016),
//    and the deep-water Green function follows the global approximation of
//    Hui Liang (2016-2018).  Both are free software under the GNU Lesser
//    General Public License, version 3 or later (src/FinGreen3D.f90:19-29,
//    src/InfGreen_Appr.f90:9-18); this file is a restatement of those
//    methods and is distributed under the same licence terms.  References:
//    Y. Liu et al., Energy Conversion and Management (2018); H. Liang, H. Wu
//    and F. Noblesse, Journal of Fluid Mechanics (2018).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "hamsx_internal.hpp"

namespace hamsx {

constexpr double kPi = 3.141592653589793;        // Const_mod (WavDynMods.f90:146)
constexpr double kG = 9.80665;                   // Const_mod gravity
#define HX_F(x) ((double)(x##f))                 // a Fortran REAL*4 literal widened to REAL*8

struct cd {
  double r, i;
};
__device__ __forceinline__ cd mk(double r, double i = 0.0) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.r, -a.i}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{a.r * s, a.i * s}; }
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{a.r * s, a.i * s}; }
__device__ __forceinline__ cd operator/(cd a, double s) { return cd{a.r / s, a.i / s}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {  // Smith's algorithm
  if (fabs(b.r) >= fabs(b.i)) {
    double q = b.i / b.r, d = b.r + b.i * q;
    return cd{(a.r + a.i * q) / d, (a.i - a.r * q) / d};
  }
  double q = b.r / b.i, d = b.r * q + b.i;
  return cd{(a.r * q + a.i) / d, (a.i * q - a.r) / d};
}
__device__ __forceinline__ cd ccos_(cd z) { return cd{cos(z.r) * cosh(z.i), -sin(z.r) * sinh(z.i)}; }
__device__ __forceinline__ cd csin_(cd z) { return cd{sin(z.r) * cosh(z.i), cos(z.r) * sinh(z.i)}; }
__device__ __forceinline__ double cabs_(cd z) { return hypot(z.r, z.i); }

// REAL**INTEGER in the square-and-multiply order flang emits (llvm.powi).
__device__ __forceinline__ double ipow(double a, int b) {
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b >>= 1;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Special functions (FinGrnExtSubs.f90)
//
// From here to the end of the finite-depth section, FMA contraction is off:
// the epsilon extrapolation of EIGENE / PIDCOCK and the adaptive quadrature
// of LINTON turn last-bit differences of their inputs into differences of up
// to ~1e-5 in their outputs (two FP64 builds of the reference's own Fortran,
// with and without FMA, disagree by that much), so the series terms are
// formed with the reference's rounding and operation order.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

// ERROR: FinGrnExtSubs.f90:62-93
__device__ inline double d_erf(double x) {
  const double eps = 1.0e-15;
  double x2 = x * x;
  if (fabs(x) < 3.5) {
    double er = 1.0, r = 1.0;
    for (int k = 1; k <= 50; k++) {
      r = r * x2 / (k + 0.5);
      er = er + r;
      if (fabs(r) <= fabs(er) * eps) break;
    }
    return 2.0 / sqrt(kPi) * x * exp(-x2) * er;
  }
  double er = 1.0, r = 1.0;
  for (int k = 1; k <= 12; k++) {
    r = -r * (k - 0.5) / x2;
    er = er + r;
  }
  double e = 1.0 - exp(-x2) / (fabs(x) * sqrt(kPi)) * er;
  return x < 0.0 ? -e : e;
}
__device__ __forceinline__ double d_erfcc(double x) { return 1.0 - d_erf(x); }
__device__ __forceinline__ double d_derfcc(double x) { return -2.0 / sqrt(kPi) * exp(-(x * x)); }

// E1XA: FinGrnExtSubs.f90:198-213
__device__ inline double d_e1xa(double x) {
  if (x == 0.0) return 1.0e300;
  if (x <= 1.0)
    return -log(x) + ((((1.07857e-3 * x - 9.76004e-3) * x + 5.519968e-2) * x - 0.24991055) * x + 0.99999193) * x -
           0.57721566;
  double es1 = (((x + 8.5733287401) * x + 18.059016973) * x + 8.6347608925) * x + 0.2677737343;
  double es2 = (((x + 9.5733223454) * x + 25.6329561486) * x + 21.0996530827) * x + 3.9584969228;
  return exp(-x) / x * es1 / es2;
}

// ENXA: FinGrnExtSubs.f90:135-157
__device__ inline double d_enxa(int n, double x) {
  double e1 = d_e1xa(x);
  if (n == 0) return exp(-x) / x;
  if (n == 1) return e1;
  double ex = exp(-x), en = e1;
  for (int k = 2; k <= n; k++) {
    en = (ex - x * e1) / (k - 1.0);
    e1 = en;
  }
  return en;
}

// EIX: FinGrnExtSubs.f90:254-282
__device__ inline double d_eix(double x) {
  if (x == 0.0) return -1.0e300;
  if (x <= 40.0) {
    double ei = 1.0, r = 1.0;
    for (int k = 1; k <= 100; k++) {
      r = r * k * x / ((k + 1.0) * (k + 1.0));
      ei = ei + r;
      if (fabs(r / ei) <= 1.0e-15) break;
    }
    return 0.5772156649015328 + log(x) + x * ei;
  }
  double ei = 1.0, r = 1.0;
  for (int k = 1; k <= 20; k++) {
    r = r * k / x;
    ei = ei + r;
  }
  return exp(x) / x * ei;
}

// Reciprocals of the series denominators of IK01A: 1/(k k), 1/(k (k + 1)) and
// 1/k for k = 1..50, each correctly rounded (constant-folded IEEE division).
struct Ik01Rcp {
  double kk[51], kk1[51], k[51];
};
constexpr Ik01Rcp make_ik01_rcp() {
  Ik01Rcp t{};
  for (int k = 1; k <= 50; k++) {
    t.kk[k] = 1.0 / (double)(k * k);
    t.kk1[k] = 1.0 / (double)(k * (k + 1));
    t.k[k] = 1.0 / (double)k;
  }
  return t;
}
__constant__ constexpr Ik01Rcp kIk01Rcp = make_ik01_rcp();

// a / b correctly rounded for the integer divisors of IK01A (b = k k or
// k (k + 1), k <= 50), given rb = RN(1/b): q = RN(a rb) is within an ulp of
// a / b, the remainder a - q b is exact in one FMA, and one correction step
// rounds to the IEEE quotient (Markstein's division by a reciprocal; checked
// against IEEE division for these divisors on 4e8 arguments).  Three
// dependent operations instead of the FP64 division sequence.
__device__ __forceinline__ double div_by_rcp(double a, double b, double rb) {
  const double q = a * rb;
  const double rem = fma(-q, b, a);
  return fma(rem, rb, q);
}

// IK01A (K0, K1): FinGrnExtSubs.f90:538-628.  The first convergence test of
// the K0 series reads an unset value in the reference (:608) and is false.
// The series quotients by k k, k (k + 1) and k are the reference's, bit for
// bit (div_by_rcp and RN(1/k) from the table).  The convergence tests
// |r / bi0| < 1e-15 and |(k0 - ww) / k0| < 1e-15 are evaluated as
// |r| < 1e-15 |bi0| and |k0 - ww| < 1e-15 |k0|: the decisions agree except
// when the ratio lies within about an ulp of 1e-15, where the series would
// take one term more or fewer (a term below 1e-15 of the sum).
__device__ inline void d_ik01a(double x, double& bk0, double& bk1) {
  if (x == 0.0) {
    bk0 = 1.0e300;
    bk1 = 1.0e300;
    return;
  }
  double x2 = x * x, bi0, bi1;
  if (x <= 18.0) {
    double r = 1.0;
    bi0 = 1.0;
    for (int k = 1; k <= 50; k++) {
      r = div_by_rcp(0.25 * r * x2, (double)(k * k), kIk01Rcp.kk[k]);
      bi0 = bi0 + r;
      if (fabs(r) < 1.0e-15 * fabs(bi0)) break;
    }
    r = 1.0;
    bi1 = 1.0;
    for (int k = 1; k <= 50; k++) {
      r = div_by_rcp(0.25 * r * x2, (double)(k * (k + 1)), kIk01Rcp.kk1[k]);
      bi1 = bi1 + r;
      if (fabs(r) < 1.0e-15 * fabs(bi1)) break;
    }
    bi1 = 0.5 * x * bi1;
  } else {
    const double A[12] = {0.125, 7.03125e-2, 7.32421875e-2, 1.1215209960938e-1, 2.2710800170898e-1,
                          5.7250142097473e-1, 1.7277275025845, 6.0740420012735, 2.4380529699556e01,
                          1.1001714026925e02, 5.5133589612202e02, 3.0380905109224e03};
    const double B[12] = {-0.375, -1.171875e-1, -1.025390625e-1, -1.4419555664063e-1, -2.7757644653320e-1,
                          -6.7659258842468e-1, -1.9935317337513, -6.8839142681099, -2.7248827311269e01,
                          -1.2159789187654e02, -6.0384407670507e02, -3.3022722944809e03};
    int k0 = x >= 50.0 ? 7 : (x >= 35.0 ? 9 : 12);
    double ca = exp(x) / sqrt(2.0 * kPi * x), xr = 1.0 / x;
    bi0 = 1.0;
    bi1 = 1.0;
    for (int k = 1; k <= k0; k++) {
      double xk = ipow(xr, k);
      bi0 = bi0 + A[k - 1] * xk;
      bi1 = bi1 + B[k - 1] * xk;
    }
    bi0 = ca * bi0;
    bi1 = ca * bi1;
  }
  double k0v;
  if (x <= 9.0) {
    double ct = -(log(x / 2.0) + 0.5772156649015329);
    double w0 = 0.0, r = 1.0, ww = 0.0;
    k0v = 0.0;
    for (int k = 1; k <= 50; k++) {
      w0 = w0 + kIk01Rcp.k[k];
      r = div_by_rcp(0.25 * r, (double)(k * k), kIk01Rcp.kk[k]) * x2;
      k0v = k0v + r * (w0 + ct);
      if (k > 1 && fabs(k0v - ww) < 1.0e-15 * fabs(k0v)) break;
      ww = k0v;
    }
    k0v = k0v + ct;
  } else {
    const double A1[8] = {0.125, 0.2109375, 1.0986328125, 1.1775970458984e01, 2.1461706161499e02,
                          5.9511522710323e03, 2.3347645606175e05, 1.2312234987631e07};
    double xr2 = 1.0 / x2, s = 1.0;
    for (int k = 1; k <= 8; k++) s = s + A1[k - 1] * ipow(xr2, k);
    k0v = 0.5 / x * s / bi0;
  }
  bk0 = k0v;
  bk1 = (1.0 / x - bi1 * k0v) / bi0;
}

// JY01B: FinGrnExtSubs.f90:663-705 (JY01BJ/BY0/BY1 evaluate the same formulas)
__device__ inline void d_jy01b(double x, double& bj0, double& bj1, double& by0, double& by1) {
  if (x == 0.0) {
    bj0 = 1.0;
    bj1 = 0.0;
    by0 = -1.0e300;
    by1 = -1.0e300;
    return;
  }
  if (x <= 4.0) {
    double t = x / 4.0, t2 = t * t;
    bj0 = ((((((-.5014415e-3 * t2 + .76771853e-2) * t2 - .0709253492) * t2 + .4443584263) * t2 - 1.7777560599) * t2 +
            3.9999973021) * t2 - 3.9999998721) * t2 + 1.0;
    bj1 = t * (((((((-.1289769e-3 * t2 + .22069155e-2) * t2 - .0236616773) * t2 + .1777582922) * t2 - .8888839649) * t2 +
                 2.6666660544) * t2 - 3.9999999710) * t2 + 1.9999999998);
    double y0 = (((((((-.567433e-4 * t2 + .859977e-3) * t2 - .94855882e-2) * t2 + .0772975809) * t2 - .4261737419) * t2 +
                   1.4216421221) * t2 - 2.3498519931) * t2 + HX_F(1.0766115157)) * t2 + .3674669052;
    double lg = 2.0 / kPi * log(x / 2.0);
    by0 = lg * bj0 + y0;
    double y1 = ((((((((.6535773e-3 * t2 - .0108175626) * t2 + .107657606) * t2 - .7268945577) * t2 + 3.1261399273) * t2 -
                    7.3980241381) * t2 + 6.8529236342) * t2 + .3932562018) * t2 - .6366197726) / x;
    by1 = lg * bj1 + y1;
    return;
  }
  double t = 4.0 / x, t2 = t * t, a0 = sqrt(2.0 / (kPi * x));
  double p0 = ((((-.9285e-5 * t2 + .43506e-4) * t2 - .122226e-3) * t2 + .434725e-3) * t2 - .4394275e-2) * t2 + .999999997;
  double q0 = t * (((((.8099e-5 * t2 - .35614e-4) * t2 + .85844e-4) * t2 - .218024e-3) * t2 + .1144106e-2) * t2 - .031249995);
  const double ta0 = x - .25 * kPi, s0 = sin(ta0), c0 = cos(ta0);
  bj0 = a0 * (p0 * c0 - q0 * s0);
  by0 = a0 * (p0 * s0 + q0 * c0);
  double p1 = ((((.10632e-4 * t2 - .50363e-4) * t2 + .145575e-3) * t2 - .559487e-3) * t2 + .7323931e-2) * t2 + 1.000000004;
  double q1 = t * (((((-.9173e-5 * t2 + .40658e-4) * t2 - .99941e-4) * t2 + .266891e-3) * t2 - .1601836e-2) * t2 + .093749994);
  const double ta1 = x - .75 * kPi, s1 = sin(ta1), c1 = cos(ta1);
  bj1 = a0 * (p1 * c1 - q1 * s1);
  by1 = a0 * (p1 * s1 + q1 * c1);
}

// Limes: Wynn epsilon table with the Mishonov-Penev selection of the largest
// auxiliary element, FinGrnExtSubs.f90:370-503.  s[0..n] is overwritten;
// a[] must hold n+1 entries.
__device__ inline double d_limes(int n, double* s, double* a) {
  double rl = s[n], amax = 0.0;
  for (int i = 0; i <= n; i++) a[i] = 0.0;
  for (int k = 1;; k++) {
    if (n - 2 * k + 1 < 0) break;
    for (int i = 0; i <= n - 2 * k + 1; i++) {
      double d = s[i + 1] - s[i];
      a[i] = (s[i + 1] != s[i]) ? a[i + 1] + 1.0 / d : a[i + 1];
    }
    if (n - 2 * k < 0) break;
    for (int i = 0; i <= n - 2 * k; i++) {
      double si1 = s[i + 1];
      s[i] = (a[i + 1] != a[i]) ? si1 + 1.0 / (a[i + 1] - a[i]) : si1;
      if (fabs(a[i]) > amax) {
        amax = fabs(a[i]);
        rl = s[i];
        if (si1 == s[i]) return rl;
      }
    }
  }
  return rl;
}

// 1/d as the FP64 division sequence (v_div_scale, v_rcp, two Newton steps,
// one correction, v_div_fmas, v_div_fixup) computes it when neither operand
// needs scaling: the same reciprocal, Newton steps and correction without
// the scale and fix-up steps, so the result is 1.0 / d bit for bit whenever
// 2^-990 <= |1/d| <= 2^750.  lo / hi record the smallest and largest |1/d|
// (fmin / fmax ignore the NaN that d = 0 gives in a lane whose quotient is
// not used), so the caller can tell afterwards that every quotient it used
// was in that range.
__device__ __forceinline__ double rcp_fast(double d, double& lo, double& hi) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(e, r, r);
  lo = __builtin_fmin(lo, fabs(r));
  hi = __builtin_fmax(hi, fabs(r));
  return r;
}

// The same Limes with s[] and a[] held in registers: the two sweeps of each
// column pair are fused into one pass (a_new[i+1] is formed just before
// s_new[i] needs it, which reads only entries the unfused order has not yet
// overwritten), and the pass is unrolled over the fixed array length in
// straight-line chunks of kLimesChunk steps, so every index is static and the
// divisions of neighbouring steps can overlap.  A chunk that straddles the
// live length also updates entries beyond it; those entries are dead (no
// later live step reads an s entry above n - 2k or an a entry above
// n - 2k + 1), and the selection scan is predicated on the live length.  Each
// live table entry is the same expression on the same operands as in
// d_limes, and the scan visits them in the same order: the result is
// identical.  Needs n <= NMAX - 1.
constexpr int kLimesChunk = 8;
// FAST = false: every quotient by IEEE division.  FAST = true: by rcp_fast;
// ok reports whether every quotient stayed in rcp_fast's exact range (the
// result is then d_limes's bit for bit; otherwise the caller recomputes).
template <int NMAX, bool FAST>
__device__ __forceinline__ double limes_reg_t(int n, double (&s)[NMAX], double (&a)[NMAX], bool& ok) {
  double rl = 0.0, amax = 0.0, lo = 1.0, hi = 1.0;
  bool done = false;
  auto recip = [&](double d) { return FAST ? rcp_fast(d, lo, hi) : 1.0 / d; };
#pragma unroll
  for (int i = 0; i < NMAX; i++) {
    if (i == n) rl = s[i];
    a[i] = 0.0;
  }
  for (int k = 1;; k++) {
    const int ls = n - 2 * k;  // last s index of this column pair; the a sweep ends at ls + 1
    if (ls + 1 < 0) break;
    {
      const double r = recip(s[1] - s[0]);
      a[0] = (s[1] != s[0]) ? a[1] + r : a[1];
    }
    if (ls < 0) break;
#pragma unroll
    for (int c = 0; c < NMAX - 2; c += kLimesChunk) {
      if (c <= ls) {
#pragma unroll
        for (int i = c; i < c + kLimesChunk && i < NMAX - 2; i++) {
          const double r = recip(s[i + 2] - s[i + 1]);
          a[i + 1] = (s[i + 2] != s[i + 1]) ? a[i + 2] + r : a[i + 2];
          const double si1 = s[i + 1];
          const double r2 = recip(a[i + 1] - a[i]);
          s[i] = (a[i + 1] != a[i]) ? si1 + r2 : si1;
          const bool ev = !done && i <= ls && fabs(a[i]) > amax;
          amax = ev ? fabs(a[i]) : amax;
          rl = ev ? s[i] : rl;
          done = done || (ev && si1 == s[i]);
        }
        if (done) {
          ok = !FAST || (lo >= 0x1p-990 && hi <= 0x1p750);
          return rl;
        }
      }
    }
  }
  ok = !FAST || (lo >= 0x1p-990 && hi <= 0x1p750);
  return rl;
}
template <int NMAX>
__device__ __forceinline__ double d_limes_reg(int n, double (&s)[NMAX], double (&a)[NMAX]) {
  bool ok;
  return limes_reg_t<NMAX, false>(n, s, a, ok);
}

// The limit of the partial sums ser[0..nt-1] (d_limes of them): the fast
// pass, and the exact pass over the same partial sums when a quotient of the
// fast pass left rcp_fast's exact range (in practice never), or the exact
// pass alone when ieee is set.
template <int NMAX>
__device__ __forceinline__ double limes_series(int nt, const double (&ser)[NMAX], int ieee) {
  double sv[NMAX], av[NMAX];
  if (!ieee) {
#pragma unroll
    for (int i = 0; i < NMAX; i++) sv[i] = i < nt ? ser[i] : 0.0;
    bool ok;
    const double r = limes_reg_t<NMAX, true>(nt - 1, sv, av, ok);
    if (__builtin_expect(ok, 1)) return r;
  }
#pragma unroll
  for (int i = 0; i < NMAX; i++) sv[i] = i < nt ? ser[i] : 0.0;
  return d_limes_reg<NMAX>(nt - 1, sv, av);
}

// ---------------------------------------------------------------------------
// Deep water: INFGREEN3D (InfGreen_Appr.f90:104-186), Havelock global fit
// (well conditioned: contraction back on)
// ---------------------------------------------------------------------------
#pragma clang fp contract(fast)
__device__ inline double d_bj0(double xx) {  // InfGreen_Appr.f90:583-626
  if (xx <= 3.0) {
    double yy = (xx / 3.0) * (xx / 3.0);
    return 0.999999999 + (-2.249999879 + (1.265623060 + (-0.316394552 + (0.044460948 + (-0.003954479 + 0.000212950 * yy) * yy) * yy) * yy) * yy) * yy;
  }
  double yy = 3.0 / xx, y2 = yy * yy;
  double f0 = 0.79788454 + (-0.00553897 + (0.00099336 + (-0.00044346 + (0.00020445 + -0.00004959 * y2) * y2) * y2) * y2) * y2;
  double th = xx - 0.25 * kPi + (-0.04166592 + (0.00239399 + (-0.00073984 + (0.00031099 + -0.00007605 * y2) * y2) * y2) * y2) * yy;
  return f0 * cos(th) / sqrt(xx);
}
__device__ inline double d_bj1(double xx) {  // InfGreen_Appr.f90:629-672
  if (xx <= 3.0) {
    double yy = xx / 3.0, y2 = yy * yy;
    return (0.5 + (-0.562499992 + (0.210937377 + (-0.039550040 + (0.004447331 + (-0.000330547 + 0.000015525 * y2) * y2) * y2) * y2) * y2) * y2) * xx;
  }
  double yy = 3.0 / xx, y2 = yy * yy;
  double f1 = 0.79788459 + (0.01662008 + (-0.00187002 + (0.00068519 + (-0.00029440 + 0.00006952 * y2) * y2) * y2) * y2) * y2;
  double th = xx - 0.75 * kPi + (0.12499895 + (-0.00605240 + (0.00135825 + (-0.00049616 + 0.00011531 * y2) * y2) * y2) * y2) * yy;
  return f1 * cos(th) / sqrt(xx);
}
__device__ inline double d_by0_far(double xx) {  // InfGreen_Appr.f90:698-714 (x > 3)
  double u = 3.0 / xx;
  double f0 = 0.79788454 - 0.00553897 * ipow(u, 2) + 0.00099336 * ipow(u, 4) - 0.00044346 * ipow(u, 6) +
              0.00020445 * ipow(u, 8) - 0.00004959 * ipow(u, 10);
  double th = xx - 0.25 * kPi - 0.04166592 * u + 0.00239399 * ipow(u, 3) - 0.00073984 * ipow(u, 5) +
              0.00031099 * ipow(u, 7) - 0.00007605 * ipow(u, 9);
  return f0 * sin(th) / sqrt(xx);
}
__device__ inline double d_by1_far(double xx) {  // InfGreen_Appr.f90:743-759 (x > 3)
  double u = 3.0 / xx;
  double f1 = 0.79788459 + 0.01662008 * ipow(u, 2) - 0.00187002 * ipow(u, 4) + 0.00068519 * ipow(u, 6) -
              0.00029440 * ipow(u, 8) + 0.00006952 * ipow(u, 10);
  double th = xx - 3.0 * kPi / 4.0 + 0.12499895 * u - 0.00605240 * ipow(u, 3) + 0.00135825 * ipow(u, 5) -
              0.00049616 * ipow(u, 7) + 0.00011531 * ipow(u, 9);
  return f1 * sin(th) / sqrt(xx);
}
__device__ inline double d_h0(double xx) {  // StruveH0, InfGreen_Appr.f90:765-818
  if (xx <= 3.0) {
    double yy = (xx / 3.0) * (xx / 3.0);
    return (1.909859164 + (-1.909855001 + (0.687514637 + (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) * yy) * yy) * yy) * (xx / 3.0);
  }
  double yy = (3.0 / xx) * (3.0 / xx);
  double c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
  double c2 = kPi * xx * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
  return c1 / c2 + d_by0_far(xx);
}
__device__ inline double d_h1(double xx) {  // StruveH1, InfGreen_Appr.f90:821-872
  if (xx <= 3.0) {
    double yy = (xx / 3.0) * (xx / 3.0);
    return (1.909859286 + (-1.145914713 + (0.294656958 + (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) * yy) * yy) * yy) * yy;
  }
  double yy = (3.0 / xx) * (3.0 / xx);
  double c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
  double c2 = kPi * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
  return c1 / c2 + d_by1_far(xx);
}

// HavelockGF (InfGreen_Appr.f90:189-580): GF = L0 + W, GFh = Ls + Wh.
__device__ inline void d_havelock(double hh, double vv, cd& gf, cd& gfh) {
  const double gama = 0.5772156649;  // truncated value of InfGreen_Appr.f90:84
  double dd = sqrt(hh * hh + vv * vv);
  double alpha = -vv / dd, beta = hh / dd, rho = dd / (1.0 + dd), t = rho;
  double dd3 = dd * dd * dd;
  double A = 1.21 + (-13.328 + (215.896 + (-1763.96 + (8418.94 + (-24314.21 + (42002.57 + (-41592.9 + (21859.0 + -4838.6 * t) * t) * t) * t) * t) * t) * t) * t) * t;
  double B = 0.938 + (5.373 + (-67.92 + (796.534 + (-4780.77 + (17137.74 + (-36618.81 + (44894.06 + (-29030.24 + 7671.22 * t) * t) * t) * t) * t) * t) * t) * t) * t;
  double C = 1.268 + (-9.747 + (209.653 + (-1397.89 + (5155.67 + (-9844.35 + (9136.4 + -3272.62 * t) * t) * t) * t) * t) * t) * t;
  double D = 0.632 + (-40.97 + (667.16 + (-6072.07 + (31127.39 + (-96293.05 + (181856.75 + (-205690.43 + (128170.2 + -33744.6 * t) * t) * t) * t) * t) * t) * t) * t) * t;
  double dA = 2.948 + (-24.53 + (249.69 + (-754.85 + (-1187.71 + (16370.75 + (-48811.41 + (68220.87 + (-46688.0 + 12622.25 * t) * t) * t) * t) * t) * t) * t) * t) * t;
  double dB = 1.11 + (2.894 + (-76.765 + (1565.35 + (-11336.19 + (44270.15 + (-97014.11 + (118879.26 + (-76209.82 + 19923.28 * t) * t) * t) * t) * t) * t) * t) * t) * t;
  double dC = 14.19 + (-148.24 + (847.8 + (-2318.58 + (3168.35 + -1590.27 * t) * t) * t) * t) * t;
  double ev = exp(vv), omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
  double pp = (log(0.5 * (dd - vv)) + gama - 2.0 * dd * dd) * ev + dd * dd - vv;
  double rr = (1.0 - beta) * A - beta * B - alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho)) + beta * (1.0 - beta) * D;
  double l0 = 2.0 * pp / (1.0 + dd3) + 2.0 * (rho * omr3 * rr);
  double ps = (beta + hh) / (dd - vv) - 2.0 * beta + 2.0 * ev * dd - hh;
  double qs = exp(-dd) * (1.0 - beta) * (1.0 + dd / (1.0 + dd3));
  double rs = beta * dA - (1.0 - alpha) * dB + beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
  double ls = 2.0 * ps / (1.0 + dd3) - 4.0 * qs + 2.0 * (rho * omr3 * rs);
  double tp = 2.0 * kPi;
  gf = cd{l0 + tp * d_h0(hh) * ev, tp * (-d_bj0(hh)) * ev};
  gfh = cd{ls + tp * (2.0 / kPi - d_h1(hh)) * ev, tp * d_bj1(hh) * ev};
}

// Rankine terms added for FLAG = 1 (only the image 1/r1) or FLAG = 0 (1/r and 1/r1)
struct PairGeom {
  double xpq, ypq, zpq, zsrc, zfld, rr, r, r1;
};

__device__ __forceinline__ PairGeom pair_geom(double xs, double xf, double ys, double yf, double zs, double zf) {
#pragma clang fp contract(off)
  PairGeom g;
  g.xpq = xs - xf;
  g.ypq = ys - yf;
  g.zpq = zf + zs;
  g.zsrc = zs;
  g.zfld = zf;
  g.rr = sqrt(g.xpq * g.xpq + g.ypq * g.ypq);
  g.r = sqrt(g.rr * g.rr + (zf - zs) * (zf - zs));
  g.r1 = sqrt(g.rr * g.rr + g.zpq * g.zpq);
  return g;
}

// INFGREEN3D, general-V branch and the V = 0 / V = -1 limits (:130-183)
__device__ inline void d_infgreen(const PairGeom& p, double v, int flag, cd g[4]) {
  double r13 = p.r1 * p.r1 * p.r1, r3 = p.r * p.r * p.r;
  for (int k = 0; k < 4; k++) g[k] = mk(0.0);
  double sgn = 1.0;
  if (fabs(v) < HX_F(1.0e-8)) {
  } else if (fabs(v + 1.0) < HX_F(1.0e-8)) {
    sgn = -1.0;
  } else {
    cd gf, gfh;
    d_havelock(v * p.rr, v * p.zpq, gf, gfh);
    double v2 = v * v;
    if (p.rr > HX_F(1.0e-6)) {
      g[1] = (-v2) * gfh * p.xpq / p.rr;
      g[2] = (-v2) * gfh * p.ypq / p.rr;
    }
    g[0] = (-v) * gf;
    g[3] = (-v2) * gf + mk(2.0 * v / p.r1);
  }
  g[0].r += sgn * (1.0 / p.r1);
  g[1].r -= sgn * (p.xpq / r13);
  g[2].r -= sgn * (p.ypq / r13);
  g[3].r -= sgn * (p.zpq / r13);
  if (flag == 0) {
    g[0].r += 1.0 / p.r;
    g[1].r -= p.xpq / r3;
    g[2].r -= p.ypq / r3;
    g[3].r -= (p.zsrc - p.zfld) / r3;
  }
}

// ---------------------------------------------------------------------------
// Finite depth: FINGREEN3D (FinGreen3D.f90:81-226) and its four regions
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
struct FreqParams {
  double v, wk, w1, h;
  const double* wvn;  // HAMSX_NK roots, wvn[0] = wk
  // Per-frequency tables (mode_tables below): every factor of a series term
  // that depends only on the frequency, the mode and one panel's z.
  const double *pm, *ac, *ut;  // [kModes] per mode
  double nm0;                  // N_0 of the propagating mode
  const double *cz, *sz;       // [kModes][n]: cos, sin of k_m (z + h)
  const double *ct, *st;       // [kModes][n]: cos, sin of u_m (z + h), u_m = m pi / h
  const double *chz, *shz;     // [n]: cosh, sinh of k (z + h)
  int n;                       // table row stride (panels)
  int js, jf;                  // table columns of zf (source panel) and zp (field panel)
  int far32 = 0;               // EIGEN region: evanescent series in FP32 (mixed precision)
  int limes_ieee = 0;          // epsilon tables by IEEE division only (diagnostic)
};

// The table entries, each formed by the expression the reference evaluates
// in place (EIGEN :300-337, EIGENE :433-474, PIDCOCK :598-640), so a table
// lookup returns the value the in-place evaluation would.
__device__ __forceinline__ void mode_scalars(int mode, const double* wvn, double v, double h, double& pm, double& ac,
                                             double& ut) {
  const double km = wvn[mode];
  const double twokmh = 2.0 * km * h;
  pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
  ac = (km * km + v * v) / ((km * km + v * v) * h - v);
  const double mp = mode * kPi;
  ut = mp / h;
}
__device__ __forceinline__ void mode_panel(int mode, const double* wvn, double h, double z, double& cz, double& sz,
                                           double& ct, double& st) {
  const double km = wvn[mode];
  const double kz = km * (z + h);
  cz = cos(kz);
  sz = sin(kz);
  const double mp = mode * kPi, umt = mp / h;
  ct = cos(umt * (z + h));
  st = sin(umt * (z + h));
}
__device__ __forceinline__ double wave_nm0(double wk, double h) {
  return h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
}

// Propagating mode (EIGEN :300-313, EIGENE :433-446, PIDCOCK :598-606),
// from J0, J1, Y0, Y1 of k R (which do not depend on the orientation)
__device__ __forceinline__ void wave_mode_from(double j0, double j1, double y0, double y1, const FreqParams& fp,
                                               bool pidcock, cd fg[3]) {
  double wk = fp.wk;
  double nm0 = fp.nm0;
  double chf = fp.chz[fp.js], chp = fp.chz[fp.jf], shf = fp.shz[fp.js];
  double p0 = kPi / nm0 * chf * chp;
  double dp0 = pidcock ? wk * kPi / nm0 * shf * chp : wk * (kPi / nm0 * shf * chp);
  fg[0] = cd{-p0 * y0, p0 * j0};
  fg[1] = cd{wk * p0 * y1, -(wk * p0) * j1};
  fg[2] = cd{-dp0 * y0, dp0 * j0};
}
__device__ inline void d_wave_mode(double r, const FreqParams& fp, bool pidcock, cd fg[3]) {
  double j0, j1, y0, y1;
  d_jy01b(fp.wk * r, j0, j1, y0, y1);
  wave_mode_from(j0, j1, y0, y1, fp, pidcock, fg);
}

// Evanescent term m (0-based mode index) of EIGEN / EIGENE (:321-337), from
// K0, K1 of k_m R.  The product cos k_m(z+h) cos k_m(zeta+h) is formed first:
// it is then the same for both orientations of a pair, bit for bit, so the
// value and d/dR series of the two orientations coincide (the reference's
// left-to-right product differs from it by rounding only).
__device__ __forceinline__ void eigen_term_from(double k0, double k1, int mode, const FreqParams& fp, double rm[3]) {
  const double km = fp.wvn[mode], pm = fp.pm[mode];
  const size_t row = (size_t)mode * fp.n;
  const double ckzf = fp.cz[row + fp.js], ckzp = fp.cz[row + fp.jf], skzf = fp.sz[row + fp.js];
  const double cc = ckzf * ckzp;
  rm[0] = pm * k0 * cc;
  rm[1] = -(pm * km * k1 * cc);
  rm[2] = -(pm * km * k0 * skzf * ckzp);
}
__device__ __forceinline__ void d_eigen_term(double r, int mode, const FreqParams& fp, double rm[3]) {
  double k0, k1;
  d_ik01a(fp.wvn[mode] * r, k0, k1);
  eigen_term_from(k0, k1, mode, fp, rm);
}

// Rankine-term removal shared by EIGEN, EIGENE and PIDCOCK
__device__ __forceinline__ void d_remove_rankine(double r, double zf, double zp, cd fg[3]) {
  double sr = sqrt(r * r + (zf - zp) * (zf - zp)), sr1 = sqrt(r * r + (zf + zp) * (zf + zp));
  fg[0].r -= 1.0 / sr + 1.0 / sr1;
  fg[1].r += (r / sr) / (sr * sr) + (r / sr1) / (sr1 * sr1);
  fg[2].r += ((zf - zp) / sr) / (sr * sr) + ((zf + zp) / sr1) / (sr1 * sr1);
}

// EIGEN: R/h >= 0.5, at most 10 terms (:281-362)
__device__ inline void d_eigen(double r, double zf, double zp, const FreqParams& fp, cd fg[3]) {
  d_wave_mode(r, fp, false, fg);
  for (int i = 2; i <= 10; i++) {
    double rm[3];
    d_eigen_term(r, i - 1, fp, rm);
    fg[0].r += rm[0];
    fg[1].r += rm[1];
    fg[2].r += rm[2];
    if (fabs(rm[0]) < HX_F(1.0e-6) && fabs(rm[1]) < HX_F(1.0e-6) && fabs(rm[2]) < HX_F(1.0e-6)) break;
  }
  d_remove_rankine(r, zf, zp, fg);
}

constexpr int kEigeneMax = 51;    // NT <= INT(-88.89*0.05 + 54.45) = 50
constexpr int kPidcockMax = 100;  // NT <= INT(-1010.10*0.0005 + 100.50) = 99

// EIGENE: 0.05 <= R/h < 0.5, epsilon-accelerated (:413-506).  The three
// partial-sum series are written once to a per-lane array and then
// extrapolated one at a time by the register-resident Limes.
__device__ __forceinline__ void d_eigene(double r, double zf, double zp, const FreqParams& fp, cd fg[3]) {
  double ser[3][kEigeneMax];
  int nt;
  {
#pragma clang fp contract(off)
    nt = (int)(HX_F(-88.89) * fabs(r / fp.h) + HX_F(54.45));
  }
  if (nt > kEigeneMax) nt = kEigeneMax;
  d_wave_mode(r, fp, false, fg);
  double g = 0.0, gr = 0.0, gz = 0.0;
  ser[0][0] = ser[1][0] = ser[2][0] = 0.0;
  for (int i = 2; i <= nt; i++) {
    double rm[3];
    d_eigen_term(r, i - 1, fp, rm);
    g = g + rm[0];
    gr = gr + rm[1];
    gz = gz + rm[2];
    ser[0][i - 1] = g;
    ser[1][i - 1] = gr;
    ser[2][i - 1] = gz;
  }
  double res[3];
#pragma unroll 1
  for (int q = 0; q < 3; q++) {
    res[q] = limes_series<kEigeneMax>(nt, ser[q], fp.limes_ieee);
  }
  fg[0].r += res[0];
  fg[1].r += res[1];
  fg[2].r += res[2];
  d_remove_rankine(r, zf, zp, fg);
}

// Chebyshev fit of the Rankine image series: IMGS/DGSR/DGSZ (:908-1102)
__device__ inline void d_image_sums(double r, double zf, double zp, double h, double out[3]) {
  const double amn[5][5] = {  // amn[n][m] = AMN(M,N)
      {0.0, -0.02525711, 0.00086546, -0.00004063, 0.00000193},
      {0.05051418, -0.00692380, 0.00073292, -0.00006636, 0.00000398},
      {0.00230838, -0.00097875, 0.00020597, -0.00003333, 0.00000524},
      {0.00012934, -0.00010879, 0.00003965, -0.00000891, 0.0},
      {0.00000913, -0.00001270, 0.00000466, 0.0, 0.0}};
  double t1 = r / h, t2 = (zf - zp) / h, t3 = (zf + zp + 2.0 * h) / h;
  double s1 = 0.0, s1r = 0.0, s1z = 0.0;
  for (int l = -1; l <= 1; l++) {
    double za = t2 + 2.0 * l, zb = t3 + 2.0 * l;
    double ra = sqrt(t1 * t1 + za * za), rb = sqrt(t1 * t1 + zb * zb);
    double ra3 = ra * ra * ra, rb3 = rb * rb * rb;
    s1 += 1.0 / ra + 1.0 / rb;
    s1r += t1 / ra3 + t1 / rb3;
    s1z += za / ra3 + zb / rb3;
  }
  // powers t^(2k) and t^(2k-1) for k = 0..4 (pow of integral exponents)
  double p1[9], p2[9], p3[9];
  p1[0] = p2[0] = p3[0] = 1.0;
  for (int k = 1; k < 9; k++) {
    p1[k] = pow(t1, (double)k);
    p2[k] = pow(t2, (double)k);
    p3[k] = pow(t3, (double)k);
  }
  double s2 = 0.0, s2r = 0.0, s2z = 0.0;
  for (int m = 0; m <= 4; m++)
    for (int n = 0; n <= 4; n++) {
      double a = amn[n][m];
      s2 += a * p1[2 * m] * (p2[2 * n] + p3[2 * n]);
      if (m >= 1) s2r += 2.0 * m * a * p1[2 * m - 1] * (p2[2 * n] + p3[2 * n]);
      if (n >= 1) s2z += 2.0 * n * a * p1[2 * m] * (p2[2 * n - 1] + p3[2 * n - 1]);
    }
  out[0] = (s1 - 2.0 + s2) / h;
  out[1] = (-s1r + s2r) / (h * h);
  out[2] = (-s1z + s2z) / (h * h);
}

// One PIDCOCK series term (:611-640) added to the partial sums, from K0, K1
// of k_m R and of u_m R.  As in eigen_term_from the cosine product is formed
// first (AC (cos cos) rather than (AC cos) cos), so the value and d/dR
// series do not depend on the orientation.
__device__ __forceinline__ void pidcock_term(double k0, double k1, double k0t, double k1t, int mode,
                                             const FreqParams& fp, double& g, double& gr, double& gz) {
  const double h = fp.h;
  const size_t row = (size_t)mode * fp.n;
  double umt = fp.ut[mode], um = fp.wvn[mode];
  double ac = fp.ac[mode];
  const double cf = fp.cz[row + fp.js], cp = fp.cz[row + fp.jf], sf = fp.sz[row + fp.js];
  const double tcf = fp.ct[row + fp.js], tcp = fp.ct[row + fp.jf], tsf = fp.st[row + fp.js];
  double am = ac * (cf * cp);
  double amt = tcf * tcp / h;
  double dam = -(um * ac * sf * cp);
  double damt = -(umt * tsf * tcp / h);
  g = g + (am * k0 - amt * k0t);
  gr = gr + (-(um * am * k1) + umt * amt * k1t);
  gz = gz + (dam * k0 - damt * k0t);
}

// PIDCOCK: 0.0005 <= R/h < 0.05 (:558-660).  Partial sums go to a per-lane
// array once; the three epsilon tables then run in registers.
__device__ __forceinline__ void d_pidcock(double r, double zf, double zp, const FreqParams& fp, cd fg[3]) {
  double ser[3][kPidcockMax];
  const double ga = 0.5772156649015328;
  double h = fp.h;
  int nt;
  {
#pragma clang fp contract(off)
    nt = (int)(HX_F(-1010.10) * fabs(r / h) + HX_F(100.50));
  }
  if (nt > kPidcockMax) nt = kPidcockMax;
  d_wave_mode(r, fp, true, fg);
  fg[0].r += 2.0 / h * (ga + log(r / 4.0 / h));
  fg[1].r += 2.0 / (r * h);
  double g = 0.0, gr = 0.0, gz = 0.0;
  ser[0][0] = ser[1][0] = ser[2][0] = 0.0;
  for (int i = 2; i <= nt; i++) {
    const int mode = i - 1;
    double k0, k1, k0t, k1t;
    d_ik01a(fp.wvn[mode] * r, k0, k1);
    d_ik01a(fp.ut[mode] * r, k0t, k1t);
    pidcock_term(k0, k1, k0t, k1t, mode, fp, g, gr, gz);
    ser[0][i - 1] = g;
    ser[1][i - 1] = gr;
    ser[2][i - 1] = gz;
  }
  double res[3];
#pragma unroll 1
  for (int q = 0; q < 3; q++) {
    res[q] = 4.0 * limes_series<kPidcockMax>(nt, ser[q], fp.limes_ieee);
  }
  double im[3];
  d_image_sums(r, zf, zp, h, im);
  fg[0].r = fg[0].r + res[0] + im[0];
  fg[1].r = fg[1].r + res[1] + im[1];
  fg[2].r = fg[2].r + res[2] + im[2];
  d_remove_rankine(r, zf, zp, fg);
}

// ---- Both orientations of one hull-panel pair ----
// For panels a and b, the pairs (source b, field a) -- orientation A -- and
// (source a, field b) -- orientation B -- have the same R, so PIDCOCK,
// EIGENE and EIGEN share every Bessel value between them: J0, J1, Y0, Y1 of
// k R and K0, K1 of k_m R (and of u_m R).  These routines evaluate those once
// per mode and form each orientation's terms, partial sums, epsilon tables
// and Rankine corrections with the expressions of the one-orientation
// routines above, so each orientation's result is the one those give, bit
// for bit.  zs / zd are orientation A's source and field z (orientation B
// swaps them); fpA and fpB differ only in their table columns (js, jf).

// EIGEN, both orientations: each keeps the reference's own break test (:330).
__device__ inline void d_eigen2(double r, double zs, double zd, const FreqParams& fpA, const FreqParams& fpB, cd fgA[3],
                                cd fgB[3]) {
  double j0, j1, y0, y1;
  d_jy01b(fpA.wk * r, j0, j1, y0, y1);
  wave_mode_from(j0, j1, y0, y1, fpA, false, fgA);
  wave_mode_from(j0, j1, y0, y1, fpB, false, fgB);
  bool goA = true, goB = true;
  for (int i = 2; i <= 10; i++) {
    double k0, k1, rm[3];
    d_ik01a(fpA.wvn[i - 1] * r, k0, k1);
    if (goA) {
      eigen_term_from(k0, k1, i - 1, fpA, rm);
      fgA[0].r += rm[0];
      fgA[1].r += rm[1];
      fgA[2].r += rm[2];
      if (fabs(rm[0]) < HX_F(1.0e-6) && fabs(rm[1]) < HX_F(1.0e-6) && fabs(rm[2]) < HX_F(1.0e-6)) goA = false;
    }
    if (goB) {
      eigen_term_from(k0, k1, i - 1, fpB, rm);
      fgB[0].r += rm[0];
      fgB[1].r += rm[1];
      fgB[2].r += rm[2];
      if (fabs(rm[0]) < HX_F(1.0e-6) && fabs(rm[1]) < HX_F(1.0e-6) && fabs(rm[2]) < HX_F(1.0e-6)) goB = false;
    }
    if (!goA && !goB) break;
  }
  d_remove_rankine(r, zs, zd, fgA);
  d_remove_rankine(r, zd, zs, fgB);
}

// ---- EIGEN region in FP32 (mixed precision; R/h >= 0.5) ----
// K0 and K1 of x > 0 in single precision by the polynomial approximations of
// Abramowitz & Stegun 9.8.1, 9.8.3, 9.8.5-9.8.8 (absolute or relative error
// below 2.2e-7 on their ranges) -- well conditioned in FP32, unlike the
// power series IK01A sums up to x = 9, whose terms cancel to 1e-8 there.
__device__ __forceinline__ void f_k01(float x, float& k0, float& k1) {
  if (x <= 2.0f) {
    const float t = x * (1.0f / 3.75f), t2 = t * t;
    const float i0 = 1.0f + t2 * (3.5156229f + t2 * (3.0899424f + t2 * (1.2067492f + t2 * (0.2659732f + t2 * (0.0360768f + t2 * 0.0045813f)))));
    const float i1 = x * (0.5f + t2 * (0.87890594f + t2 * (0.51498869f + t2 * (0.15084934f + t2 * (0.02658733f + t2 * (0.00301532f + t2 * 0.00032411f))))));
    const float y = 0.25f * x * x, lg = logf(0.5f * x);
    k0 = -lg * i0 + (-0.57721566f + y * (0.42278420f + y * (0.23069756f + y * (0.03488590f + y * (0.00262698f + y * (0.00010750f + y * 0.0000074f))))));
    k1 = lg * i1 + (1.0f + y * (0.15443144f + y * (-0.67278579f + y * (-0.18156897f + y * (-0.01919402f + y * (-0.00110404f + y * -0.00004686f)))))) / x;
  } else {
    const float y = 2.0f / x, e = expf(-x) * rsqrtf(x);
    k0 = e * (1.25331414f + y * (-0.07832358f + y * (0.02189568f + y * (-0.01062446f + y * (0.00587872f + y * (-0.00251540f + y * 0.00053208f))))));
    k1 = e * (1.25331414f + y * (0.23498619f + y * (-0.03655620f + y * (0.01504268f + y * (-0.00780353f + y * (0.00325614f + y * -0.00068245f))))));
  }
}

// The evanescent modes of EIGEN (:321-337) for both orientations in FP32,
// each with the reference's break test (:330); the propagating mode (FP64)
// is in fgA / fgB already and the modes are added to it.  The Rankine terms
// are not removed here: the caller adds the remainder that the removal and
// FINGREEN3D's re-addition leave (green_rankine_far).
__device__ inline void f_eigen_modes2(double r, const FreqParams& fpA, const FreqParams& fpB, cd fgA[3], cd fgB[3]) {
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f;
  bool goA = true, goB = true;
  const float rf = (float)r;
  for (int i = 2; i <= 10; i++) {
    const int mode = i - 1;
    const float km = (float)fpA.wvn[mode], pm = (float)fpA.pm[mode];
    float k0, k1;
    f_k01(km * rf, k0, k1);
    const size_t row = (size_t)mode * fpA.n;
    const float cA = (float)fpA.cz[row + fpA.js], cB = (float)fpA.cz[row + fpA.jf];
    const float cc = cA * cB, pk0 = pm * k0, pk1 = pm * km * k1;
    if (goA) {
      const float r0 = pk0 * cc, r1 = -(pk1 * cc), r2 = -(pm * km * k0 * (float)fpA.sz[row + fpA.js] * cB);
      a0 += r0;
      a1 += r1;
      a2 += r2;
      if (fabsf(r0) < 1.0e-6f && fabsf(r1) < 1.0e-6f && fabsf(r2) < 1.0e-6f) goA = false;
    }
    if (goB) {
      const float r0 = pk0 * cc, r1 = -(pk1 * cc), r2 = -(pm * km * k0 * (float)fpB.sz[row + fpB.js] * cA);
      b0 += r0;
      b1 += r1;
      b2 += r2;
      if (fabsf(r0) < 1.0e-6f && fabsf(r1) < 1.0e-6f && fabsf(r2) < 1.0e-6f) goB = false;
    }
    if (!goA && !goB) break;
  }
  fgA[0].r += a0;
  fgA[1].r += a1;
  fgA[2].r += a2;
  fgB[0].r += b0;
  fgB[1].r += b1;
  fgB[2].r += b2;
}

// EIGEN with FP32 evanescent modes, both orientations (no Rankine removal).
__device__ inline void d_eigen2_far32(double r, const FreqParams& fpA, const FreqParams& fpB, cd fgA[3], cd fgB[3]) {
  double j0, j1, y0, y1;
  d_jy01b(fpA.wk * r, j0, j1, y0, y1);
  wave_mode_from(j0, j1, y0, y1, fpA, false, fgA);
  wave_mode_from(j0, j1, y0, y1, fpB, false, fgB);
  f_eigen_modes2(r, fpA, fpB, fgA, fgB);
}

// EIGENE, both orientations: the value and d/dR series are shared (see
// eigen_term_from), so four partial-sum series and four epsilon tables.
__device__ __forceinline__ void d_eigene2(double r, double zs, double zd, const FreqParams& fpA,
                                          const FreqParams& fpB, cd fgA[3], cd fgB[3]) {
  double ser[4][kEigeneMax];
  int nt;
  {
#pragma clang fp contract(off)
    nt = (int)(HX_F(-88.89) * fabs(r / fpA.h) + HX_F(54.45));
  }
  if (nt > kEigeneMax) nt = kEigeneMax;
  double j0, j1, y0, y1;
  d_jy01b(fpA.wk * r, j0, j1, y0, y1);
  wave_mode_from(j0, j1, y0, y1, fpA, false, fgA);
  wave_mode_from(j0, j1, y0, y1, fpB, false, fgB);
  double g[4] = {0.0, 0.0, 0.0, 0.0};  // value, d/dR, d/dz (A), d/dz (B)
#pragma unroll
  for (int q = 0; q < 4; q++) ser[q][0] = 0.0;
  for (int i = 2; i <= nt; i++) {
    double k0, k1, rm[3];
    d_ik01a(fpA.wvn[i - 1] * r, k0, k1);
    eigen_term_from(k0, k1, i - 1, fpA, rm);
    g[0] = g[0] + rm[0];
    g[1] = g[1] + rm[1];
    g[2] = g[2] + rm[2];
    eigen_term_from(k0, k1, i - 1, fpB, rm);
    g[3] = g[3] + rm[2];
#pragma unroll
    for (int q = 0; q < 4; q++) ser[q][i - 1] = g[q];
  }
  double res[4];
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    res[q] = limes_series<kEigeneMax>(nt, ser[q], fpA.limes_ieee);
  }
  fgA[0].r += res[0];
  fgA[1].r += res[1];
  fgA[2].r += res[2];
  fgB[0].r += res[0];
  fgB[1].r += res[1];
  fgB[2].r += res[3];
  d_remove_rankine(r, zs, zd, fgA);
  d_remove_rankine(r, zd, zs, fgB);
}

// PIDCOCK, both orientations: K0, K1 of k_m R and u_m R once per mode; the
// value and d/dR series are shared (see pidcock_term), four epsilon tables.
__device__ __forceinline__ void d_pidcock2(double r, double zs, double zd, const FreqParams& fpA,
                                           const FreqParams& fpB, cd fgA[3], cd fgB[3]) {
  double ser[4][kPidcockMax];
  const double ga = 0.5772156649015328;
  double h = fpA.h;
  int nt;
  {
#pragma clang fp contract(off)
    nt = (int)(HX_F(-1010.10) * fabs(r / h) + HX_F(100.50));
  }
  if (nt > kPidcockMax) nt = kPidcockMax;
  double j0, j1, y0, y1;
  d_jy01b(fpA.wk * r, j0, j1, y0, y1);
  wave_mode_from(j0, j1, y0, y1, fpA, true, fgA);
  wave_mode_from(j0, j1, y0, y1, fpB, true, fgB);
  const double lg = 2.0 / h * (ga + log(r / 4.0 / h)), rg = 2.0 / (r * h);
  fgA[0].r += lg;
  fgA[1].r += rg;
  fgB[0].r += lg;
  fgB[1].r += rg;
  double g[4] = {0.0, 0.0, 0.0, 0.0};  // value, d/dR, d/dz (A), d/dz (B)
  double gb0 = 0.0, gb1 = 0.0;         // orientation B's copies of the shared sums
#pragma unroll
  for (int q = 0; q < 4; q++) ser[q][0] = 0.0;
  for (int i = 2; i <= nt; i++) {
    const int mode = i - 1;
    double k0, k1, k0t, k1t;
    d_ik01a(fpA.wvn[mode] * r, k0, k1);
    d_ik01a(fpA.ut[mode] * r, k0t, k1t);
    pidcock_term(k0, k1, k0t, k1t, mode, fpA, g[0], g[1], g[2]);
    pidcock_term(k0, k1, k0t, k1t, mode, fpB, gb0, gb1, g[3]);
#pragma unroll
    for (int q = 0; q < 4; q++) ser[q][i - 1] = g[q];
  }
  double res[4];
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    res[q] = 4.0 * limes_series<kPidcockMax>(nt, ser[q], fpA.limes_ieee);
  }
  double im[3];
  d_image_sums(r, zs, zd, h, im);
  fgA[0].r = fgA[0].r + res[0] + im[0];
  fgA[1].r = fgA[1].r + res[1] + im[1];
  fgA[2].r = fgA[2].r + res[2] + im[2];
  d_image_sums(r, zd, zs, h, im);
  fgB[0].r = fgB[0].r + res[0] + im[0];
  fgB[1].r = fgB[1].r + res[1] + im[1];
  fgB[2].r = fgB[2].r + res[3] + im[2];
  d_remove_rankine(r, zs, zd, fgA);
  d_remove_rankine(r, zd, zs, fgB);
}

// ---- LINTON (R/h < 0.0005), FinGreen3D.f90:712-861 and its helpers ----

// Integrands: F1/F2 (:1904-1998) and G1/G2/G3 (:2001-2209)
struct Integrand {
  int kind;  // 0: F1/F2, 1: G1..G3
  int tag;
  int no;
  double r, wvn, v, zf, zp, h;
};

__device__ inline double d_integrand(const Integrand& f, double t) {
  if (fabs(t) < 1.0e-8) return 0.0;
  if (f.kind == 0) {
    double w2t = f.wvn * f.wvn * t;
    double e = exp(-(f.r * f.r / 4.0 / t) + (f.no == 1 ? w2t : -w2t));
    if (f.tag == 1) return e / 4.0 / kPi / t;
    return -(f.r * e / 8.0 / kPi / (t * t));
  }
  const double val[4] = {-f.zp - f.zf, 2.0 * f.h - f.zp + f.zf, 2.0 * f.h + f.zp - f.zf, 4.0 * f.h + f.zp + f.zf};
  const double dval[4] = {-1.0, 1.0, -1.0, 1.0};
  double v = f.v, e = exp(v * v * (t * t) - f.r * f.r / 4.0 / (t * t)), sum = 0.0;
  for (int i = 0; i < 4; i++) {
    double ev = exp(-v * val[i]), arg = val[i] / 2.0 / t - v * t;
    if (f.tag == 3)
      sum = sum - v * dval[i] * ev * d_erfcc(arg) + ev * d_derfcc(arg) * dval[i] / 2.0 / t;
    else
      sum = sum + ev * d_erfcc(arg);
  }
  if (f.tag == 2) return f.r / 2.0 / (t * t * t) * e * sum;
  return e * sum / t;
}

// AQUAD / AQUADF (:2387-2452, :2515-2532, :2599-2665, :2728-2745): 7-point
// Gauss / 15-point Kronrod with the reference's bisection rule, run
// depth-first on a bounded stack; leaves are summed left to right.
constexpr int kQuadStack = 64;
__device__ __noinline__ double d_aquad(const Integrand& f, double a0, double b0, double eps, int* overflow) {
  const double xgq[7] = {-0.949107912342759, -0.741531185599394, -0.405845151377397, 0.0,
                         0.405845151377397, 0.741531185599394, 0.949107912342759};
  const double wgq[7] = {0.129484966168870, 0.279705391489277, 0.381830050505119, 0.417959183673469,
                         0.381830050505119, 0.279705391489277, 0.129484966168870};
  const double xkq[8] = {-0.991455371120813, -0.864864423359769, -0.586087235467691, -0.207784955007898,
                         0.207784955007898, 0.586087235467691, 0.864864423359769, 0.991455371120813};
  const double wkq_odd[8] = {0.022935322010529, 0.104790010322250, 0.169004726639267, 0.204432940075298,
                             0.204432940075298, 0.169004726639267, 0.104790010322250, 0.022935322010529};
  const double wkq_even[7] = {0.063092092629979, 0.140653259715525, 0.190350578064785, 0.209482141084728,
                              0.190350578064785, 0.140653259715525, 0.063092092629979};
  double sa[kQuadStack], sb[kQuadStack];
  int sp = 1;
  sa[0] = a0;
  sb[0] = b0;
  double total = 0.0;
  while (sp > 0) {
    --sp;
    double a = sa[sp], b = sb[sp], hw = (b - a) / 2.0;
    double gi = 0.0, ki = 0.0, stf[7];
    for (int i = 0; i < 7; i++) {
      stf[i] = d_integrand(f, 0.5 * (b + a + (b - a) * xgq[i]));
      gi += wgq[i] * (b - a) / 2.0 * stf[i];
    }
    // Kronrod sum in the reference's order: odd nodes are new, even reuse stf
    for (int i = 0; i < 15; i++) {
      if (i & 1)
        ki += wkq_even[i >> 1] * (b - a) / 2.0 * stf[i >> 1];
      else
        ki += wkq_odd[i >> 1] * (b - a) / 2.0 * d_integrand(f, 0.5 * (b + a + (b - a) * xkq[i >> 1]));
    }
    (void)hw;
    double err = pow(200.0 * fabs(ki - gi), 1.5);
    if (err > eps && fabs(a - b) > eps) {
      if (sp + 2 > kQuadStack) {
        *overflow = 1;
        total += ki;
        continue;
      }
      double m = (a + b) / 2.0;
      sa[sp] = m;  // right half below left half: left is taken first
      sb[sp] = b;
      ++sp;
      sa[sp] = a;
      sb[sp] = m;
      ++sp;
    } else {
      total += ki;
    }
  }
  return total;
}

// SA0 / DSA0 / SAM / DSAM series (:1389-1694)
__device__ inline double d_sa0(double r, double pa, double wk, bool deriv) {
  const double ga = 0.5772156649015328, eps = HX_F(1.0e-6);
  double rm = 100.0, s = deriv ? 0.0 : -ga - 2.0 * log(wk * pa / 2.0), sgn = -1.0, facn = 1.0;
  double lg = log(wk * r / 2.0), rp = r / pa, wp = wk * pa / 2.0, wr = wk * r / 2.0, psi = 0.0;
  for (int n = 1; fabs(rm) > eps; n++) {
    psi = psi + 1.0 / n;
    double f2 = facn * facn;
    if (!deriv) {
      double en = d_enxa(n + 1, rp * rp);
      rm = sgn * pow(rp, 2.0 * n) / facn / n - pow(wp, 2.0 * n) / facn * en -
           2.0 * sgn / f2 * pow(wr, 2.0 * n) * (lg + ga - psi);
    } else {
      double en = d_enxa(n, rp * rp);
      rm = 2.0 * sgn * pow(rp, 2.0 * n - 1.0) / facn / pa + pow(wp, 2.0 * n) / facn * en * 2.0 * r / (pa * pa) -
           2.0 * sgn * n * wk / f2 * pow(wr, 2.0 * n - 1.0) * (lg + ga - psi) - 2.0 * sgn / f2 * pow(wr, 2.0 * n) / r;
    }
    s = s + rm;
    sgn = -sgn;
    facn = facn * (n + 1);
    if (n > 400) break;
  }
  return s / 4.0 / kPi;
}

__device__ inline double d_sam(double r, double pa, double wvn, bool deriv) {
  const double eps = HX_F(1.0e-6);
  double rm = 100.0, s = 0.0, sgn = deriv ? -1.0 : 1.0, facn = 1.0, rp = r / pa;
  double x = (wvn * pa / 2.0) * (wvn * pa / 2.0);
  for (int n = deriv ? 1 : 0; fabs(rm) > eps; n++) {
    double en = d_enxa(n + 1, x);
    if (!deriv)
      rm = sgn * pow(rp, 2.0 * n) / facn * en;
    else
      rm = 2.0 * n * sgn * pow(rp, 2.0 * n - 1.0) / facn / pa * en;
    s = s + rm;
    sgn = -sgn;
    facn = facn * (n + 1);
    if (n > 400) break;
  }
  return s / 4.0 / kPi;
}

// COEF / DCOEF coefficient m (1-based) of the LINTON eigen-expansion (:1168-1340)
__device__ inline double d_coef(bool deriv, int m, double r, double pa, double a, const FreqParams& fp, int* ovf) {
  const double eps = HX_F(1.0e-6), step = 100.0;
  double dnt = pa * pa / 4.0;
  if (m == 1) {
    double wk = fp.wk;
    if (fabs(r) < HX_F(1.0e-6)) return deriv ? 0.0 : -d_eix(wk * wk * dnt) / 4.0 / kPi;
    if (r / a <= 0.5) return d_sa0(r, pa, wk, deriv);
    double j0, j1, y0, y1;
    d_jy01b(wk * r, j0, j1, y0, y1);
    Integrand f{0, deriv ? 2 : 1, 1, r, wk, 0, 0, 0, 0};
    double q = d_aquad(f, 0.0, dnt, eps, ovf);
    return deriv ? wk * y1 / 4.0 - q : -y0 / 4.0 - q;
  }
  double km = fp.wvn[m - 1];
  if (fabs(r) < HX_F(1.0e-6)) return deriv ? 0.0 : d_e1xa(km * km * dnt) / 4.0 / kPi;
  if (r / a <= 1.0) return d_sam(r, pa, km, deriv);
  Integrand f{0, deriv ? 2 : 1, m, r, km, 0, 0, 0, 0};
  double p1 = dnt + step;
  for (int it = 0; it < 100000 && fabs(d_integrand(f, p1)) > eps; it++) p1 = p1 + step;
  return d_aquad(f, dnt, p1, eps, ovf);
}

__device__ inline cd d_nm(double h, cd km) {  // NM, FinGreen3D.f90:2285-2295
  cd t = cd{2.0 * km.r * h, 2.0 * km.i * h};
  cd q = cdiv(csin_(t), t);
  return cd{h / 2.0 * (1.0 + q.r), h / 2.0 * q.i};
}

__device__ __noinline__ void d_linton(double r, double zf, double zp, const FreqParams& fp, cd fg[3], int* ovf) {
  const double eps = HX_F(1.0e-6);
  double v = fp.v, wk = fp.wk, h = fp.h;
  double a;
  if (v < 0.02) a = 3324.2 * (v * v * v) - 258.89 * (v * v) + 8.3958 * v + 1.3208;
  else if (v < 0.1) a = -30.303 * (v * v * v) - 2.8138 * (v * v) + 2.7740 * v + 0.05052;
  else if (v < 1.0) a = 1.5699 * (v * v * v) - 2.5075 * (v * v) + 1.8316 * v + 0.1077;
  else a = -0.000069641 * (v * v * v) - 0.0017082 * (v * v) + 0.45566 * v + 0.80551;
  a = fabs(a) / (wk * h);
  double pa = a * h;
  double sr = sqrt(r * r + (zf - zp) * (zf - zp));
  double sr1 = sqrt(r * r + (2.0 * h + zf + zp) * (2.0 * h + zf + zp));
  double sr2 = sqrt(r * r + (zf + zp) * (zf + zp));

  // evanescent-mode part (section III)
  cd g0 = mk(0.0), g1 = mk(0.0), g2 = mk(0.0);
  for (int i = 1; i <= 20; i++) {
    double cof = d_coef(false, i, r, pa, a, fp, ovf);
    double dcof = d_coef(true, i, r, pa, a, fp, ovf);
    cd km = (i == 1) ? cd{0.0, -wk} : cd{fp.wvn[i - 1], 0.0};
    cd nm = d_nm(h, km);
    cd cf = ccos_(km * (zf + h)), cp = ccos_(km * (zp + h)), sf = csin_(km * (zf + h));
    cd cc = cf * cp;
    cd rm1 = -(cdiv(mk(cof), nm) * cc);
    cd rm2 = -(cdiv(mk(dcof), nm) * cc);
    cd rm3 = cdiv(cof * km, nm) * sf * cp;
    g0 = g0 + rm1;
    g1 = g1 + rm2;
    g2 = g2 + rm3;
    if (i > 2 && cabs_(rm1) < eps && cabs_(rm2) < eps && cabs_(rm3) < eps) break;
  }

  // error-function part (section IV)
  const double val[4] = {-zp - zf, 2.0 * h - zp + zf, 2.0 * h + zp - zf, 4.0 * h + zp + zf};
  double e1 = 0.0, e2 = 0.0, e3 = 0.0, fp4 = 4.0 * kPi;
  for (int i = 0; i < 4; i++) {
    double ro = sqrt(r * r + val[i] * val[i]);
    double dror = r / ro, droz = val[i] * ((i & 1) ? 1.0 : -1.0) / ro;
    double q = ro / pa, ec = d_erfcc(q), dc = d_derfcc(q);
    e1 += ec / (fp4 * ro);
    e2 += (dc / pa * dror * ro - dror * ec) / (fp4 * ro * ro);
    e3 += (dc / pa * droz * ro - droz * ec) / (fp4 * ro * ro);
  }
  double dsr1r = r / sr1, dsr1z = (2.0 * h + zf + zp) / sr1, dsr2r = r / sr2, dsr2z = (zf + zp) / sr2;
  double ec1 = d_erfcc(sr1 / pa), dc1 = d_derfcc(sr1 / pa);
  if (fabs(sr) > HX_F(1.0e-6)) {
    double ec = d_erfcc(sr / pa), dc = d_derfcc(sr / pa), dsrr = r / sr, dsrz = (zf - zp) / sr;
    e1 = e1 + (ec - 1.0) / (fp4 * sr) + ec1 / (fp4 * sr1) - 1.0 / (fp4 * sr2);
    e2 = e2 - (ec - 1.0) / (fp4 * sr * sr) * dsrr + dc / (fp4 * sr * pa) * dsrr - ec1 / (fp4 * sr1 * sr1) * dsr1r +
         dc1 / (fp4 * sr1 * pa) * dsr1r + dsr2r / (fp4 * sr2 * sr2);
    e3 = e3 - (ec - 1.0) / (fp4 * sr * sr) * dsrz + dc / (fp4 * sr * pa) * dsrz - ec1 / (fp4 * sr1 * sr1) * dsr1z +
         dc1 / (fp4 * sr1 * pa) * dsr1z + dsr2z / (fp4 * sr2 * sr2);
  } else {
    e1 = e1 + ec1 / (fp4 * sr1) - 1.0 / (fp4 * sr2);
    e2 = e2 - ec1 / (fp4 * sr1 * sr1) * dsr1r + dc1 / (fp4 * sr1 * pa) * dsr1r + dsr2r / (fp4 * sr2 * sr2);
    e3 = e3 - ec1 / (fp4 * sr1 * sr1) * dsr1z + dc1 / (fp4 * sr1 * pa) * dsr1z + dsr2z / (fp4 * sr2 * sr2);
  }

  // integral part (section V) and the propagating mode
  double c = v / 2.0 / kPi;
  Integrand q1{1, 1, 0, r, 0.0, v, zf, zp, h}, q2{1, 2, 0, r, 0.0, v, zf, zp, h}, q3{1, 3, 0, r, 0.0, v, zf, zp, h};
  g0 = g0 - mk(e1) - mk(c * d_aquad(q1, 0.0, pa / 2.0, eps, ovf));
  g1 = g1 - mk(e2) + mk(c * d_aquad(q2, 0.0, pa / 2.0, eps, ovf));
  g2 = g2 - mk(e3) - mk(c * d_aquad(q3, 0.0, pa / 2.0, eps, ovf));
  double bj0, bj1, by0, by1;
  d_jy01b(wk * r, bj0, bj1, by0, by1);
  cd nm = d_nm(h, cd{0.0, -wk});
  double chf = cosh(wk * (zf + h)), chp = cosh(wk * (zp + h)), shf = sinh(wk * (zf + h));
  cd ci4 = cdiv(cd{0.0, 0.25}, nm);  // CI/4/NM
  g0 = g0 - ci4 * (bj0 * chf * chp);
  g1 = g1 + (ci4 * wk) * (bj1 * chf * chp);
  g2 = g2 - (ci4 * wk) * (bj0 * shf * chp);
  fg[0] = (-fp4) * g0;
  fg[1] = (-fp4) * g1;
  fg[2] = (-fp4) * g2;
}

// Regions of FINGREEN3D_R (:183-191) by R/h, and the binning key the
// assembly sorts pairs by: region, then the series length NT of EIGENE /
// PIDCOCK, so that the lanes of a wave run the same code for the same number
// of terms.  The key is a function of the geometry and h only.
__device__ __forceinline__ int fin_region(double rr, double h) {
  double rhp = fabs(rr / h);
  if (rhp < 0.0005) return kRegLinton;
  if (rhp < 0.05) return kRegPidcock;
  if (rhp < 0.5) return kRegEigene;
  return kRegEigen;
}

// Sub-key (see hamsx_internal.hpp), clamped so that any NT maps inside its
// region's range.
__device__ __forceinline__ int fin_subkey(double rr, double h) {
  const int reg = fin_region(rr, h);
  int nt;
  switch (reg) {
    case kRegLinton:
      return 0;
    case kRegPidcock: {
#pragma clang fp contract(off)
      nt = (int)(HX_F(-1010.10) * fabs(rr / h) + HX_F(100.50));
      return kSubPidcock + min(kSubEigene - kSubPidcock - 1, max(0, 99 - nt));
    }
    case kRegEigene: {
#pragma clang fp contract(off)
      nt = (int)(HX_F(-88.89) * fabs(rr / h) + HX_F(54.45));
      return kSubEigene + min(kSubEigen - kSubEigene - 1, max(0, 50 - nt));
    }
    default:
      return kSubEigen;
  }
}
// FINGREEN3D (:81-155): the Green function and its gradient from the
// region routine's fg (value, d/dR, d/dz) ...
__device__ __forceinline__ void green_from_fg(const PairGeom& p, const cd fg[3], cd g[4]) {
  if (p.rr > HX_F(1.0e-6)) {
    g[1] = fg[1] * p.xpq / p.rr;
    g[2] = fg[1] * p.ypq / p.rr;
  }
  g[0] = fg[0];
  g[3] = fg[2];
}
// ... plus the Rankine terms of the image (and of the source when FLAG = 0)
__device__ __forceinline__ void green_rankine(const PairGeom& p, double sgn, int flag, cd g[4]) {
  double r13 = p.r1 * p.r1 * p.r1, r3 = p.r * p.r * p.r;
  g[0].r += sgn * (1.0 / p.r1);
  g[1].r -= sgn * (p.xpq / r13);
  g[2].r -= sgn * (p.ypq / r13);
  g[3].r -= sgn * (p.zpq / r13);
  if (flag == 0) {
    g[0].r += 1.0 / p.r;
    g[1].r -= p.xpq / r3;
    g[2].r -= p.ypq / r3;
    g[3].r -= (p.zsrc - p.zfld) / r3;
  }
}

// What the EIGEN routine's Rankine removal (:344-357) and green_rankine
// leave together, exactly: the removed 1/r and 1/r1 terms equal the added
// ones (sr = r, sr1 = r1), so nothing for FLAG = 0 and the source term's
// negative for FLAG = 1 (the Newman integrals carry 1/r there).  The FP32
// path adds this instead of subtracting and re-adding the terms.
__device__ __forceinline__ void green_rankine_far(const PairGeom& p, int flag, cd g[4]) {
  if (flag == 0) return;
  const double r3 = p.r * p.r * p.r;
  g[0].r -= 1.0 / p.r;
  g[1].r += p.xpq / r3;
  g[2].r += p.ypq / r3;
  g[3].r += (p.zsrc - p.zfld) / r3;
}

// FINGREEN3D with FINGREEN3D_R's region split (:183-191).  REG fixes the
// region at compile time for pairs binned by fin_region; kRegAny decides per
// pair.
template <int REG>
__device__ inline void d_fingreen_r(const PairGeom& p, const FreqParams& fp, int flag, cd g[4], int* ovf) {
  for (int k = 0; k < 4; k++) g[k] = mk(0.0);
  double sgn = 1.0;
  if (fabs(fp.v) < HX_F(1.0e-8)) {
  } else if (fabs(fp.v + 1.0) < HX_F(1.0e-8)) {
    sgn = -1.0;
  } else {
    cd fg[3];
    const int reg = REG == kRegAny ? fin_region(p.rr, fp.h) : REG;
    if (reg == kRegEigen && fp.far32) {
      cd fgB[3];
      d_eigen2_far32(p.rr, fp, fp, fg, fgB);
      green_from_fg(p, fg, g);
      green_rankine_far(p, flag, g);
      return;
    }
    if (reg == kRegLinton)
      d_linton(p.rr, p.zsrc, p.zfld, fp, fg, ovf);
    else if (reg == kRegPidcock)
      d_pidcock(p.rr, p.zsrc, p.zfld, fp, fg);
    else if (reg == kRegEigene)
      d_eigene(p.rr, p.zsrc, p.zfld, fp, fg);
    else
      d_eigen(p.rr, p.zsrc, p.zfld, fp, fg);
    green_from_fg(p, fg, g);
  }
  green_rankine(p, sgn, flag, g);
}

// Both orientations of a hull-panel pair in region REG (PIDCOCK, EIGENE or
// EIGEN): pA / pB and fpA / fpB as for d_eigene2.  Each orientation's g is
// what d_fingreen_r<REG> gives for it.
template <int REG>
__device__ inline void d_fingreen2(const PairGeom& pA, const PairGeom& pB, const FreqParams& fpA,
                                   const FreqParams& fpB, int flagA, int flagB, cd gA[4], cd gB[4]) {
  static_assert(REG == kRegPidcock || REG == kRegEigene || REG == kRegEigen, "paired regions");
  for (int k = 0; k < 4; k++) gA[k] = gB[k] = mk(0.0);
  double sgn = 1.0;
  if (fabs(fpA.v) < HX_F(1.0e-8)) {
  } else if (fabs(fpA.v + 1.0) < HX_F(1.0e-8)) {
    sgn = -1.0;
  } else {
    cd fgA[3], fgB[3];
    if (REG == kRegEigen && fpA.far32) {
      d_eigen2_far32(pA.rr, fpA, fpB, fgA, fgB);
      green_from_fg(pA, fgA, gA);
      green_from_fg(pB, fgB, gB);
      green_rankine_far(pA, flagA, gA);
      green_rankine_far(pB, flagB, gB);
      return;
    }
    if (REG == kRegPidcock)
      d_pidcock2(pA.rr, pA.zsrc, pA.zfld, fpA, fpB, fgA, fgB);
    else if (REG == kRegEigene)
      d_eigene2(pA.rr, pA.zsrc, pA.zfld, fpA, fpB, fgA, fgB);
    else
      d_eigen2(pA.rr, pA.zsrc, pA.zfld, fpA, fpB, fgA, fgB);
    green_from_fg(pA, fgA, gA);
    green_from_fg(pB, fgB, gB);
  }
  green_rankine(pA, sgn, flagA, gA);
  green_rankine(pB, sgn, flagB, gB);
}

__device__ inline void d_fingreen(const PairGeom& p, const FreqParams& fp, int flag, cd g[4], int* ovf) {
  d_fingreen_r<kRegAny>(p, fp, flag, g, ovf);
}

#pragma clang fp contract(fast)
// ---------------------------------------------------------------------------
// Newman (1986) panel integral, SGLINTBD_MULTI_TRI/QUAD (SingularIntgrMulti.f90:44-445).
// The per-panel local frame (origin, unit vectors, projected vertices) is
// frequency- and field-point-independent and is precomputed on the host.
// ---------------------------------------------------------------------------
constexpr int kFrameLen = 24;  // po[3] ui[3] uj[3] uk[3] ksi[5] eta[5] pad[2]

__device__ inline void d_newman(const double* fr, int ncn, double xf, double yf, double zf, double& sij, double dij[3]) {
  double px = xf - fr[0], py = yf - fr[1], pz = zf - fr[2];
  double X = fr[3] * px + fr[4] * py + fr[5] * pz;
  double Y = fr[6] * px + fr[7] * py + fr[8] * pz;
  double Z = fr[9] * px + fr[10] * py + fr[11] * pz;
  const double* ksi = fr + 12;
  const double* eta = fr + 17;
  double st = 0.0, sxt = 0.0, syt = 0.0, szt = 0.0, z2 = Z * Z;
  bool quad = (ncn == 4);
  for (int i = 0; i < ncn; i++) {
    double k0 = ksi[i], k1 = ksi[i + 1], e0 = eta[i], e1 = eta[i + 1];
    double dksi = k1 - k0, deta = e1 - e0;
    double r1 = sqrt((k0 - X) * (k0 - X) + (e0 - Y) * (e0 - Y) + z2);
    double r2 = sqrt((k1 - X) * (k1 - X) + (e1 - Y) * (e1 - Y) + z2);
    double l = sqrt(dksi * dksi + deta * deta);
    if (quad && fabs(l) < 1.0e-8) continue;
    if (!(fabs(Z) < 1.0e-6)) {
      double b1 = deta * ((k0 - X) * (k0 - X) + z2) - dksi * (k0 - X) * (e0 - Y);
      double a1 = r1 * Z * dksi;
      double b2 = deta * ((k1 - X) * (k1 - X) + z2) - dksi * (k1 - X) * (e1 - Y);
      double a2 = r2 * Z * dksi;
      szt += atan2(b1 * a2 - b2 * a1, a1 * a2 + b1 * b2);
    }
    if (fabs(r1 + r2 - l) < 1.0e-6) continue;
    double lg = log((r1 + r2 + l) / (r1 + r2 - l));
    sxt -= deta / l * lg;
    syt += dksi / l * lg;
    double t = (deta * (X - k0) - dksi * (Y - e0)) / l * lg;
    st = quad ? st - t : st + t;
  }
  double sgn = quad ? -1.0 : 1.0;
  sij = quad ? st + Z * szt : st - Z * szt;
  dij[0] = sgn * (fr[3] * sxt + fr[6] * syt + fr[9] * szt);
  dij[1] = sgn * (fr[4] * sxt + fr[7] * syt + fr[10] * szt);
  dij[2] = sgn * (fr[5] * sxt + fr[8] * syt + fr[11] * szt);
}

}  // namespace hamsx
