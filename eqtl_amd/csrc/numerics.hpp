// Scalar numerics shared by the HIP kernels and their host-side unit hooks:
//  * Student-t two-sided tail (util.cpp:209-264 p_from_t/betai/betacf),
//  * the ACAT tan term and final p exactly as the reference evaluates them in
//    x87 80-bit long double (stats_extra.cpp:16-31), emulated with 64/128-bit
//    integer arithmetic so that saturation (p < 2^-66 → tan of the rounded
//    pi/2, a large *negative* number) and rounding match bit for bit,
//  * double-double accumulation for the per-gene tan sums.
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define EQ_HD __host__ __device__ inline
#else
#define EQ_HD inline  // host-only build of the same scalar code (unit checks)
#endif

namespace eqtl_amd {

// ---------------------------------------------------------------------------
// Student t tail
// ---------------------------------------------------------------------------
EQ_HD double betacf_dev(double a, double b, double x) {
  const double eps = 3e-12, fpmin = 1e-30;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (fabs(d) < fpmin) d = fpmin;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 200; ++m) {
    const int m2 = 2 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < eps) break;
  }
  return h;
}

// lbeta = lgamma(df/2) + lgamma(1/2) - lgamma(df/2 + 1/2), precomputed per df.
EQ_HD double p_from_t_dev(double t, double df, double lbeta) {
  if (!isfinite(t) || df <= 0) return 1.0;
  const double x = df / (df + t * t);
  const double a = 0.5 * df, b = 0.5;
  double p;
  if (x < 0 || x > 1) return 1.0;
  if (x == 0 || x == 1) {
    p = x;
  } else {
    const double front = exp(log(x) * a + log(1 - x) * b - lbeta) / a;
    if (x < (a + 1) / (a + b + 2)) p = front * betacf_dev(a, b, x);
    else p = 1.0 - front * betacf_dev(b, a, 1 - x) * a / b;
  }
  return fmin(1.0, fmax(0.0, p));
}

// ---------------------------------------------------------------------------
// double-double
// ---------------------------------------------------------------------------
struct dd {
  double hi, lo;
};
EQ_HD dd dd_add(dd a, dd b) {
  double s = a.hi + b.hi;
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  double hi = s + e;
  return dd{hi, e - (hi - s)};
}
EQ_HD dd dd_add_d(dd a, double b) { return dd_add(a, dd{b, 0.0}); }

// ---------------------------------------------------------------------------
// x87 long double emulation of tan((0.5L - p) * kPi), kPi = 3.14159...846L
// (stats_extra.cpp:20). Values are M * 2^E with a 64-bit mantissa M.
// ---------------------------------------------------------------------------
typedef unsigned __int128 u128;

// kPi as an 80-bit long double: 64-bit mantissa of pi rounded to nearest.
constexpr uint64_t kPiMant = 0xC90FDAA22168C235ull;  // kPi = kPiMant * 2^-62
// pi/2 to 128 bits: floor(pi * 2^126) (value = kHalfPi128 * 2^-127).
EQ_HD u128 half_pi_128() {
  return (static_cast<u128>(0xC90FDAA22168C234ull) << 64) | static_cast<u128>(0xC4C6628B80DC1CD1ull);
}

// Round a positive 128-bit integer to 64 significant bits (ties to even).
// Returns the mantissa in [2^63, 2^64); *shift gets the binary exponent change
// (value = mantissa * 2^(*shift)).
EQ_HD uint64_t rne64(u128 v, int* shift) {
  const uint64_t vh = static_cast<uint64_t>(v >> 64), vl = static_cast<uint64_t>(v);
  const int lead = vh ? 127 - __builtin_clzll(vh) : 63 - __builtin_clzll(vl);
  if (lead <= 63) {
    *shift -= (63 - lead);
    return vl << (63 - lead);
  }
  const int s = lead - 63;
  u128 q = v >> s;
  const u128 rem = v & ((static_cast<u128>(1) << s) - 1);
  const u128 half = static_cast<u128>(1) << (s - 1);
  if (rem > half || (rem == half && (q & 1))) q += 1;
  int sh = s;
  if (q >> 64) {  // rounding carried into bit 64
    q >>= 1;
    sh += 1;
  }
  *shift += sh;
  return static_cast<uint64_t>(q);
}

EQ_HD double u128_to_double(u128 v) {
  return static_cast<double>(static_cast<uint64_t>(v >> 64)) * 18446744073709551616.0 +
         static_cast<double>(static_cast<uint64_t>(v));
}

// tan((0.5L - p) * kPi) evaluated as x87 would, for 0 < p < 1 (finite).
EQ_HD double acat_term_ld(double p) {
  if (p == 0.5) return 0.0;
  // step 1: x = RN64(0.5 - p) = xm * 2^xe with sign xs
  uint64_t xm;
  int xe;
  bool neg = false;
  if (p >= 0.25) {
    const double x = 0.5 - p;  // exact (Sterbenz / p in [0.5,1))
    neg = x < 0;
    int e;
    const double f = frexp(fabs(x), &e);   // f in [0.5,1)
    xm = static_cast<uint64_t>(ldexp(f, 64));  // exact: 53 significant bits
    xe = e - 64;
  } else if (p <= 0x1p-66) {
    xm = 1ull << 63;  // 0.5 exactly
    xe = -64;
  } else {
    int e;
    const double f = frexp(p, &e);
    const uint64_t mp = static_cast<uint64_t>(ldexp(f, 53));
    const int ep = e - 53;                // p = mp * 2^ep, ep >= -118
    const u128 X = (static_cast<u128>(1) << 117) - (static_cast<u128>(mp) << (ep + 118));
    int sh = 0;
    xm = rne64(X, &sh);
    xe = sh - 118;
  }
  // step 2: y = RN64(x * kPi)
  int sh = 0;
  const uint64_t ym = rne64(static_cast<u128>(xm) * static_cast<u128>(kPiMant), &sh);
  const int ye = xe - 62 + sh;  // |y| = ym * 2^ye
  // step 3: tan(y) (tanl is correctly rounded to ~1 ulp of long double)
  if (p >= 0.25 && p <= 0.75) {
    const double yd = ldexp(static_cast<double>(ym), ye);  // |y| <= ~pi/4
    const double t = tan(yd);
    return neg ? -t : t;
  }
  // near pi/2: tan(|y|) = cot(z), z = pi/2 - |y| (exact in 128-bit fixed point)
  const int s = ye + 127;  // |y| * 2^127 = ym << s ; |y| in (pi/4, pi/2+) → s in {63, 64}
  const u128 yfix = static_cast<u128>(ym) << s;
  const u128 hp = half_pi_128();
  const double z = (hp >= yfix) ? ldexp(u128_to_double(hp - yfix), -127)
                                : -ldexp(u128_to_double(yfix - hp), -127);
  const double t = cos(z) / sin(z);
  return neg ? -t : t;
}

// The same term in double precision for 1e-6 <= p <= 1 - 1e-6, from
// sin/cos of pi times an exactly formed argument (tan((0.5 - p) pi) =
// cot(pi p) = tan(pi (0.5 - p)) = -cot(pi (1 - p)) on the three ranges). There
// the reference's own rounding of its long double argument moves the term by
// at most ~5e-14 relative (|dy| <= 2^-64 next to pi/2, divided by pi min(p, 1-p)),
// so this agrees with the x87 result to that level; outside the range (tiny p,
// or p within 1e-6 of 1) the x87 emulation above is used.
EQ_HD double acat_term_fast(double p) {
#ifdef __HIP_DEVICE_COMPILE__
  // one sincospi on the reduced argument, no branches (the lanes of a wave
  // hold unrelated p): p < 1/4 -> cot(pi p), p <= 3/4 -> tan(pi (1/2 - p)),
  // p > 3/4 -> -cot(pi (1 - p))
  const bool lo = p < 0.25, hi = p > 0.75;
  const double r = lo ? p : (hi ? 1.0 - p : 0.5 - p);
  double s, c;
  sincospi(r, &s, &c);
  const bool cot = lo || hi;
  const double q = (cot ? c : s) / (cot ? s : c);
  return hi ? -q : q;
#else
  const long double pi = 3.14159265358979323846264338327950288L;
  if (p < 0.25) return static_cast<double>(cosl(pi * p) / sinl(pi * p));
  if (p <= 0.75) return static_cast<double>(tanl(pi * (0.5L - p)));
  return static_cast<double>(-cosl(pi * (1.0L - p)) / sinl(pi * (1.0L - p)));
#endif
}

EQ_HD double acat_term(double p) {
  return (p >= 1e-6 && p <= 1.0 - 1e-6) ? acat_term_fast(p) : acat_term_ld(p);
}

// AcatAcc::add skip rules (stats_extra.cpp:17-19); returns false if skipped.
EQ_HD bool acat_accept(double p, double* term) {
  if (!isfinite(p) || p < 0.0) return false;
  if (p >= 1.0 - 1e-12) return false;
  if (p <= 0.0) p = 1e-300;
  *term = acat_term(p);
  return true;
}

// AcatAcc::p(): 0.5 - (double)(atanl(T)/kPi), T = sum/used, clamped to [0,1].
// (double)(...) rounds a value near 0.5 to the double grid, which quantises
// tiny results to multiples of 2^-54; reproduced by forming v in double.
EQ_HD double acat_p_final(dd sum, int used) {
  if (used == 0) return 1.0;
  const double T = (sum.hi + sum.lo) / static_cast<double>(used);
  const double kPiD = 3.141592653589793238462643383279502884;
  double v;
  if (fabs(T) > 1.0) {
    const double small = atan(1.0 / fabs(T)) / kPiD;  // atan(|T|) = pi/2 - atan(1/|T|)
    v = 0.5 - small;
    if (T < 0) v = -v;
  } else {
    v = atan(T) / kPiD;
  }
  double ac = 0.5 - v;
  if (ac < 0) ac = 0;
  if (ac > 1) ac = 1;
  return ac;
}

}  // namespace eqtl_amd

namespace eqtl_amd {

// ---------------------------------------------------------------------------
// Tabulated two-sided t tail for one df (all tests of a scan share df).
// The reference evaluates p from x = df/(df + t^2) (util.cpp:259-264):
//   y = 1 - x <= 1 - x_sw : p = 1 - s,  s = front(x) betacf(b, a, y) a/b
//   otherwise            : p = front(x) betacf(a, b, x)
// with a = df/2, b = 1/2, x_sw = (a+1)/(a+b+2). Both pieces are smooth in the
// variables below, so each is tabulated as per-interval degree-7 polynomials
// fitted at Chebyshev nodes of the exact expression, and evaluated at the same
// rounded x the reference forms (so its rounding of x carries over):
//   A: r(u) = s/u with u = sqrt(y), 64 uniform intervals on [0, sqrt(1-x_sw)]
//   M: ln p as a function of u, 64 uniform intervals from sqrt(1-x_sw) to
//      u_m (|t| = 8, at most u = 0.9): one exp and no log for the |t| most
//      tests have, which would otherwise take piece B's two logs
//   B: ln p as a function of w = ln(-a ln x), uniform intervals up to
//      ln p = -760 (p = 0 beyond: the reference's exp() underflows there).
// ---------------------------------------------------------------------------
constexpr int kTailDeg = 8;      // coefficients per interval
constexpr int kTailA = 64;       // intervals in piece A
constexpr int kTailM = 64;       // intervals in piece M
constexpr int kTailB = 512;      // intervals in piece B

struct TTail {
  double df, half_df;
  double u_sw, inv_h_a;          // piece A: u range end, 1 / interval width
  double w_lo, w_hi, inv_hw_b;   // piece B: w range, 1 / interval width
  double x_zero;                 // x at/below which p == 0
  double coef_a[kTailA * kTailDeg];
  double coef_b[kTailB * kTailDeg];
  double u_m, inv_h_m;           // piece M: u range end, 1 / interval width (0: no piece M)
  double coef_m[kTailM * kTailDeg];
};
static_assert(offsetof(TTail, coef_m) == offsetof(TTail, u_m) + 2 * sizeof(double), "TTail piece M layout");
static_assert(sizeof(double) * 8 == offsetof(TTail, coef_a), "TTail header layout");

EQ_HD double tail_poly(const double* c, double z) {
  double r = c[kTailDeg - 1];
  for (int k = kTailDeg - 2; k >= 0; --k) r = r * z + c[k];
  return r;
}

// The scalar header (8 doubles) and piece A's coefficients lead the struct, so
// a kernel can stage those 8 + kTailA * kTailDeg doubles in LDS and read only
// piece B (larger |t|) from global memory.
constexpr int kTailHead = 8 + kTailA * kTailDeg;

// LDS copy of the header and pieces A and M for kernels that evaluate many
// tails: each interval's coefficients kTailLdsStride doubles apart (one more
// than the coefficients, so the lanes' data-dependent intervals fall on
// different LDS banks). head: kTailLdsHead doubles, mid: kTailLdsMid.
constexpr int kTailLdsStride = kTailDeg + 1;
constexpr int kTailLdsHead = 8 + kTailA * kTailLdsStride;
constexpr int kTailLdsMid = 2 + kTailM * kTailLdsStride;

// p at x = df / (df + t^2) and y = 1 - x (y >= 0).
// head: df, half_df, u_sw, inv_h_a, w_lo, w_hi, inv_hw_b, x_zero, coef_a[];
// mid: u_m, inv_h_m, coef_m[]; coef_b: piece B. sa / sm: doubles from one
// interval's coefficients of piece A / M to the next (kTailDeg in TTail; a
// padded copy in LDS spreads the intervals over the banks).
EQ_HD double p_tail_xy(const double* head, const double* mid, const double* coef_b, double x, double y,
                       int sa = kTailDeg, int sm = kTailDeg) {
  const double u = sqrt(y);
  if (u <= head[2]) {
    const double q = u * head[3];
    int i = static_cast<int>(q);
    if (i >= kTailA) i = kTailA - 1;
    const double z = 2.0 * (q - i) - 1.0;
    return fmin(1.0, fmax(0.0, 1.0 - u * tail_poly(head + 8 + i * sa, z)));
  }
  if (u <= mid[0]) {
    const double q = (u - head[2]) * mid[1];
    int i = static_cast<int>(q);
    if (i >= kTailM) i = kTailM - 1;
    const double z = 2.0 * (q - i) - 1.0;
    return fmin(1.0, fmax(0.0, exp(tail_poly(mid + 2 + i * sm, z))));
  }
  if (x <= head[7]) return 0.0;
  const double w = log(-head[1] * log(x));
  const double q = (w - head[4]) * head[6];
  int i = static_cast<int>(q);
  if (i < 0) i = 0;
  if (i >= kTailB) i = kTailB - 1;
  const double z = 2.0 * (q - i) - 1.0;
  return fmin(1.0, fmax(0.0, exp(tail_poly(coef_b + i * kTailDeg, z))));
}

// p from t, with x formed as the reference forms it (util.cpp:259-264).
EQ_HD double p_from_t_tab2(const double* head, const double* mid, const double* coef_b, double t,
                           int sa = kTailDeg, int sm = kTailDeg) {
  const double df = head[0];
  if (!isfinite(t) || df <= 0) return 1.0;
  const double x = df / (df + t * t);
  return p_tail_xy(head, mid, coef_b, x, 1.0 - x, sa, sm);
}

EQ_HD double p_from_t_tab(const TTail& T, double t) { return p_from_t_tab2(&T.df, &T.u_m, T.coef_b, t); }

// Fills the padded LDS copy (head, mid) with the nthreads threads of a
// workgroup (thread tid); the caller synchronises before use.
EQ_HD void tail_to_lds(const TTail& T, double* head, double* mid, int tid, int nthreads) {
  for (int i = tid; i < 8; i += nthreads) head[i] = (&T.df)[i];
  for (int i = tid; i < kTailA * kTailDeg; i += nthreads)
    head[8 + (i / kTailDeg) * kTailLdsStride + i % kTailDeg] = T.coef_a[i];
  for (int i = tid; i < 2; i += nthreads) mid[i] = (&T.u_m)[i];
  for (int i = tid; i < kTailM * kTailDeg; i += nthreads)
    mid[2 + (i / kTailDeg) * kTailLdsStride + i % kTailDeg] = T.coef_m[i];
}

// Host: fit one interval (Chebyshev interpolation, then monomial in z).
inline void tail_fit_interval(double (*f)(double, const void*), const void* ctx, double lo, double hi,
                              double* out) {
  const int N = kTailDeg;
  double fv[kTailDeg], c[kTailDeg];
  const double PI = 3.14159265358979323846;
  for (int i = 0; i < N; ++i) {
    const double zi = cos(PI * (i + 0.5) / N);
    fv[i] = f(0.5 * (lo + hi) + 0.5 * (hi - lo) * zi, ctx);
  }
  for (int k = 0; k < N; ++k) {
    double s = 0;
    for (int i = 0; i < N; ++i) s += fv[i] * cos(PI * k * (i + 0.5) / N);
    c[k] = (k == 0 ? 1.0 : 2.0) * s / N;
  }
  double Tm[kTailDeg][kTailDeg] = {};
  Tm[0][0] = 1.0;
  if (N > 1) Tm[1][1] = 1.0;
  for (int k = 2; k < N; ++k)
    for (int j = 0; j < N; ++j) Tm[k][j] = (j > 0 ? 2.0 * Tm[k - 1][j - 1] : 0.0) - Tm[k - 2][j];
  for (int j = 0; j < N; ++j) {
    double s = 0;
    for (int k = 0; k < N; ++k) s += c[k] * Tm[k][j];
    out[j] = s;
  }
}

struct TailCtx {
  double df, lbeta, a, b;
};

// piece A: s/u at u = sqrt(1 - x), s = front * betacf(b, a, y) * a / b
inline double tail_fa(double u, const void* vp) {
  const TailCtx& c = *static_cast<const TailCtx*>(vp);
  if (u < 1e-9) u = 1e-9;
  const double y = u * u, x = 1.0 - y;
  const double front = exp(log1p(-y) * c.a + log(y) * c.b - c.lbeta) / c.a;
  return front * betacf_dev(c.b, c.a, y) * c.a / c.b / u;
}

// piece B: ln p at x(w), -a ln x = e^w
inline double tail_fb(double w, const void* vp) {
  const TailCtx& c = *static_cast<const TailCtx*>(vp);
  const double lnx = -exp(w) / c.a;
  const double x = exp(lnx);
  const double lfront = lnx * c.a + log1p(-x) * c.b - c.lbeta - log(c.a);
  return lfront + log(betacf_dev(c.a, c.b, x));
}

// piece M: ln p at u = sqrt(1 - x)
inline double tail_fm(double u, const void* vp) {
  const TailCtx& c = *static_cast<const TailCtx*>(vp);
  const double y = u * u, x = 1.0 - y;
  const double lfront = log1p(-y) * c.a + log(y) * c.b - c.lbeta - log(c.a);
  return lfront + log(betacf_dev(c.a, c.b, x));
}

inline void tail_build(TTail& T, double df, double lbeta) {
  T.df = df;
  T.half_df = 0.5 * df;
  const double a = 0.5 * df, b = 0.5;
  const double x_sw = (a + 1) / (a + b + 2);
  TailCtx c{df, lbeta, a, b};
  T.u_sw = sqrt(1.0 - x_sw);
  const double ha = T.u_sw / kTailA;
  T.inv_h_a = 1.0 / ha;
  for (int i = 0; i < kTailA; ++i) tail_fit_interval(tail_fa, &c, i * ha, (i + 1) * ha, T.coef_a + i * kTailDeg);
  const double v_lo = -a * log(x_sw);
  double v_hi = v_lo;
  while (tail_fb(log(v_hi), &c) > -760.0) v_hi *= 1.25;
  T.w_lo = log(v_lo);
  T.w_hi = log(v_hi);
  const double hb = (T.w_hi - T.w_lo) / kTailB;
  T.inv_hw_b = 1.0 / hb;
  for (int i = 0; i < kTailB; ++i)
    tail_fit_interval(tail_fb, &c, T.w_lo + i * hb, T.w_lo + (i + 1) * hb, T.coef_b + i * kTailDeg);
  T.x_zero = exp(-v_hi / a);
  T.u_m = fmin(sqrt(64.0 / (df + 64.0)), 0.9);
  if (T.u_m > T.u_sw) {
    const double hm = (T.u_m - T.u_sw) / kTailM;
    T.inv_h_m = 1.0 / hm;
    for (int i = 0; i < kTailM; ++i)
      tail_fit_interval(tail_fm, &c, T.u_sw + i * hm, T.u_sw + (i + 1) * hm, T.coef_m + i * kTailDeg);
  } else {
    T.u_m = T.u_sw;
    T.inv_h_m = 0.0;
  }
}

}  // namespace eqtl_amd
