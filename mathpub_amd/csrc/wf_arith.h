// wf_arith.h -- exact modular arithmetic mod N = p^2 for the base-2 Fermat quotient on
// gfx950 (CDNA4), one prime per lane.
//
// Two exact implementations of r = 2^(p-1) mod p^2 (A001220/wieferich.c:893-949 computes
// the same residue with 2 x 64-bit Montgomery limbs and a 5-bit window):
//
//  * fp64_pow2 / fp64_pow2x2: p < F64_PMAX = 9 * 2^47.  The residue is held in p-adic digits x = a + b p
//    with a, b signed integers stored exactly in doubles.  A squaring needs a^2 mod p,
//    floor(a^2/p) and 2ab mod p; each product is split exactly by FMA (h = x*y,
//    l = fma(x, y, -h)) and reduced with a rounded quotient k = rint(h / p) and an exact FMA
//    remainder.  gfx950 issues v_fma_f64 at the full VALU rate, faster than v_mad_u64_u32.
//
//  * mont_pow2<L, FULL>: any odd p < 2^64.  L 32-bit limbs, R = 2^(32 L), products
//    from v_mad_u64_u32 / v_mul_lo_u32 chains (operand scanning), lazy values in [0, 2N)
//    while N < R/4 and full reduction with the carry bit (FULL) when N may reach 2^128.
//    The base-2 ladder is left to right: square, then double when the exponent bit is 1.
//
// Both return the residue; the caller forms q = (r - 1) p^-1 mod 2^64 and the FLT check
// q < p exactly as wieferich.c:1051-1063 does.
#pragma once
#include <stdint.h>
#ifdef WF_HOST_EMU
// Host build of the same arithmetic for CPU tests (tests/emu): IEEE-754 doubles with FMA
// give the same results as the GPU's v_fma_f64 / v_rndne_f64 / v_ldexp_f64.
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
static inline uint64_t __umul64hi(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }
#define __builtin_fma(a, b, c) std::fma(a, b, c)
#define __builtin_rint(a) std::rint(a)
#define __builtin_ldexp(a, d) std::ldexp(a, d)
#define __builtin_floor(a) std::floor(a)
#else
#include <hip/hip_runtime.h>
#endif

namespace wf {

typedef uint64_t u64;
typedef uint32_t u32;

// p^-1 mod 2^64 by Newton iteration from 3p XOR 2 (correct to 5 bits), 5->80 bits.
__device__ __forceinline__ u64 inv64(u64 p) {
  u64 x = (3 * p) ^ 2;
  x *= 2 - p * x;
  x *= 2 - p * x;
  x *= 2 - p * x;
  x *= 2 - p * x;
  return x;
}

// ---------------------------------------------------------------------------------
// FP64 p-adic path, p < F64_PMAX = 9 * 2^47 (about 1.267e15)
// ---------------------------------------------------------------------------------
//
// Invariants for p < 2^50 (u = 2^-53, all values integers held exactly), for any number of
// squarings:
//   |a| <= 0.59 p (0.5 p right after a window step), |b| <= 1.9 p.
// Squaring, x^2 = a^2 + 2ab p (mod p^2), quotients rounded once inside an FMA:
//   a^2 = k1 p + a'      k1 = round(fl(a*a) / p): |k1 - a^2/p| <= 0.5 + 0.25 alpha^2,
//                       so alpha' = 0.5 + 0.25 alpha^2 (fixed point 0.59);
//   ab  = k2 p + r       k2 = round(fl(a*b) / p): |r| <= (0.5 + 0.25 alpha beta) p;
//   b'  = k1 + 2r        |b'| <= alpha^2 p + 1 + (1 + 0.5 alpha beta) p, so
//                       beta' = 1.34 + 0.29 beta (fixed point 1.9).
// All quotients stay below 2^51 in magnitude (round_mul's range) and every remainder below
// 2^52, so each FMA result is exact.  ip may be 1/p correctly rounded or within
// 2^-53 (1 + 2^-7) of it (f64_recip_near): the 0.25 alpha^2 terms grow by 2^-8 of themselves,
// and the fixed points above keep their digits.
// From 2^50 to F64_PMAX the quotient errors grow with c = p / 2^52 (0.25 alpha^2 becomes
// about c alpha^2) and the iteration has no fixed point below 2 for beta, but a ladder never
// runs more than five squarings between window steps, and a window step (f64_mul_pow2*,
// f64_mul_small*) leaves |a| <= p/2 (1 + 2^-20), |b| <= 1.1 p.  From there, at c = 0.28125:
// alpha 0.571, 0.592, 0.599, 0.601, 0.602 and beta 1.56, 1.83, 1.96, 2.02, 2.05 after one to
// five squarings, the largest rounded quotient is 0.69 * 2^51 and |b| stays below 0.29 * 2^53
// (tests/test_arith_emu.py propagates these bounds in exact rationals).
constexpr u64 F64_PMAX = 9ull << 47;
struct F64Mod {
  double p, ip;   // p, 1/p
};

// Round-to-nearest integer of x*y + 0 for |x*y| < 2^51: adding 1.5 * 2^52 inside the FMA
// leaves a sum in [2^52, 2^53), whose unit in the last place is 1, so the single rounding
// of the FMA is the rounding to an integer.  Rounding the exact product (rather than
// rint(fl(x*y)), also two instructions) keeps the quotient error bounds below.
#define WF_RND_MAGIC 6755399441055744.0
__device__ __forceinline__ double round_mul(double x, double y) {
  return __builtin_fma(x, y, WF_RND_MAGIC) - WF_RND_MAGIC;
}

__device__ __forceinline__ void f64_sqr(double& a, double& b, const F64Mod& m) {
  const double h1 = a * a;
  const double l1 = __builtin_fma(a, a, -h1);
  const double k1 = round_mul(h1, m.ip);
  const double a1 = __builtin_fma(-k1, m.p, h1) + l1;      // a^2 - k1 p
  const double h2 = a * b;
  const double l2 = __builtin_fma(a, b, -h2);
  const double k2 = round_mul(h2, m.ip);
  const double r2 = __builtin_fma(-k2, m.p, h2) + l2;      // ab - k2 p
  b = __builtin_fma(2.0, r2, k1);                          // k1 + 2ab - 2 k2 p
  a = a1;
}

// x *= 2^d for 0 <= d <= 31: a 2^d and b 2^d are exact (power-of-two scaling), each is
// reduced by an exact FMA remainder, and the carry ka moves from a into b.
__device__ __forceinline__ void f64_mul_pow2(double& a, double& b, int d, const F64Mod& m) {
  const double A = __builtin_ldexp(a, d);
  const double ka = round_mul(A, m.ip);
  a = __builtin_fma(-ka, m.p, A);
  const double B = __builtin_ldexp(b, d);
  const double kb = round_mul(B, m.ip);
  const double b1 = __builtin_fma(-kb, m.p, B) + ka;
  const double kc = round_mul(b1, m.ip);
  b = __builtin_fma(-kc, m.p, b1);
}

// The same for p >= 2^32, one reduction shorter: the carry |ka| <= 0.59 * 2^31 + 1 < 0.3 p,
// so b' = (b 2^d - kb p) + ka stays within 0.8 p, inside the squaring's bound |b| <= 1.9 p.
__device__ __forceinline__ void f64_mul_pow2_big(double& a, double& b, int d, const F64Mod& m) {
  const double A = __builtin_ldexp(a, d);
  const double ka = round_mul(A, m.ip);
  a = __builtin_fma(-ka, m.p, A);
  const double B = __builtin_ldexp(b, d);
  const double kb = round_mul(B, m.ip);
  b = __builtin_fma(-kb, m.p, B) + ka;
}

// a in [0, p), then b in [0, p): the canonical digits of r = a + b p.
__device__ __forceinline__ void f64_normalise(double& a, double& b, const F64Mod& m) {
  if (a < 0.0) { a += m.p; b -= 1.0; }
  else if (a >= m.p) { a -= m.p; b += 1.0; }
  const double kb = __builtin_floor(b * m.ip);
  b = __builtin_fma(-kb, m.p, b);
  if (b < 0.0) b += m.p;
  else if (b >= m.p) b -= m.p;
}

// r = 2^e mod p^2 as (a, b) normalised to 0 <= a, b < p (r = a + b p).  ebits = the
// number of bits to process (the wave's longest exponent), W = window width.
template <int W>
__device__ __forceinline__ void fp64_pow2(u64 p, u64 e, int ebits, double& ra, double& rb) {
  F64Mod m;
  m.p = (double)p;
  m.ip = 1.0 / m.p;
  const int nw = (ebits + W - 1) / W;
  int sh = (nw - 1) * W;
  double a = 1.0, b = 0.0;
  f64_mul_pow2(a, b, (int)(e >> sh), m);
  for (sh -= W; sh >= 0; sh -= W) {
#pragma unroll
    for (int s = 0; s < W; s++) f64_sqr(a, b, m);
    f64_mul_pow2(a, b, (int)((e >> sh) & ((1u << W) - 1)), m);
  }
  f64_normalise(a, b, m);
  ra = a;
  rb = b;
}

// Two independent ladders per lane (two primes, given as doubles pd = p exactly): the two
// instruction streams interleave, which hides FP64 latency without more wavefronts.  The
// digits are left as the last window step leaves them (|a| <= p/2 (1 + 2^-20), |b| <= p),
// for f64_quotient or f64_normalise.
//
// Leading bits.  In general x = 2^d for the leading window digit d.  With BIG (both p >=
// 2^32, ebits >= 32) the six leading bits t < 64 are taken at once: 2^t = a + k p with
// k = round(2^t / p) < 2^31 < p/2 and a = 2^t - k p exact, |a| <= p/2; then a first window
// of (ebits - 6) mod W bits, so that the W-bit windows end at bit 0.  Against the plain
// leading digit this saves 6 - (ebits mod W or W) squarings: one at 50 bits.
//
// ip0, ip1: 1/p within 2^-53 (1 + 2^-7) of it, relative (1.0 / pd, or f64_recip_near).
template <int W, bool BIG = false>
__device__ __forceinline__ void fp64_pow2x2_r(double pd0, double ip0, u64 e0, double pd1, double ip1, u64 e1,
                                              int ebits, double& a0, double& b0, double& a1, double& b1) {
  F64Mod m0, m1;
  m0.p = pd0;
  m0.ip = ip0;
  m1.p = pd1;
  m1.ip = ip1;
  int sh;
  if (BIG) {
    sh = ebits - 6;
    const double x0 = __builtin_ldexp(1.0, (int)(e0 >> sh)), x1 = __builtin_ldexp(1.0, (int)(e1 >> sh));
    b0 = round_mul(x0, m0.ip);
    a0 = __builtin_fma(-b0, m0.p, x0);
    b1 = round_mul(x1, m1.ip);
    a1 = __builtin_fma(-b1, m1.p, x1);
    const int r = sh % W;
    if (r) {
      sh -= r;
      for (int s = 0; s < r; s++) {
        f64_sqr(a0, b0, m0);
        f64_sqr(a1, b1, m1);
      }
      f64_mul_pow2_big(a0, b0, (int)((e0 >> sh) & ((1u << r) - 1)), m0);
      f64_mul_pow2_big(a1, b1, (int)((e1 >> sh) & ((1u << r) - 1)), m1);
    }
  } else {
    const int nw = (ebits + W - 1) / W;
    sh = (nw - 1) * W;
    // the leading digit: x = 2^d directly when that is already below p
    const int d0 = (int)(e0 >> sh), d1 = (int)(e1 >> sh);
    a0 = __builtin_ldexp(1.0, d0); b0 = 0.0;
    a1 = __builtin_ldexp(1.0, d1); b1 = 0.0;
    if (a0 >= pd0) { a0 = 1.0; f64_mul_pow2(a0, b0, d0, m0); }
    if (a1 >= pd1) { a1 = 1.0; f64_mul_pow2(a1, b1, d1, m1); }
  }
  for (sh -= W; sh >= 0; sh -= W) {
#pragma unroll
    for (int s = 0; s < W; s++) {
      f64_sqr(a0, b0, m0);
      f64_sqr(a1, b1, m1);
    }
    if (BIG) {   // both p >= 2^32
      f64_mul_pow2_big(a0, b0, (int)((e0 >> sh) & ((1u << W) - 1)), m0);
      f64_mul_pow2_big(a1, b1, (int)((e1 >> sh) & ((1u << W) - 1)), m1);
    } else {
      f64_mul_pow2(a0, b0, (int)((e0 >> sh) & ((1u << W) - 1)), m0);
      f64_mul_pow2(a1, b1, (int)((e1 >> sh) & ((1u << W) - 1)), m1);
    }
  }
}

template <int W, bool BIG = false>
__device__ __forceinline__ void fp64_pow2x2(double pd0, u64 e0, double pd1, u64 e1, int ebits, double& a0,
                                            double& b0, double& a1, double& b1) {
  fp64_pow2x2_r<W, BIG>(pd0, 1.0 / pd0, e0, pd1, 1.0 / pd1, e1, ebits, a0, b0, a1, b1);
}

// 1/p for p = base + f, base >= 2^40 and 0 <= f < 2^22 (pd = p, fd = f, exact as doubles), from
// ib = 1/base rounded to nearest.  The first-order estimate y = ib (1 - f ib) has a relative
// error below (f/base)^2 + 3 * 2^-53 < 2^-35, and one Newton step y + y (1 - p y) squares it:
// the result is within 2^-53 + 2^-70 of 1/p, relative (the inner FMA's own rounding adds
// 2^-35 * 2^-53).  Four FP64 instructions per prime against about twelve for the correctly
// rounded division; the squaring's bounds take 1/p with a relative error up to 2^-53 (1 + 2^-7).
__host__ __device__ __forceinline__ double f64_recip_near(double ib, double fd, double pd) {
  const double y = ib * __builtin_fma(-fd, ib, 1.0);
  const double e = __builtin_fma(-pd, y, 1.0);
  return __builtin_fma(y, e, y);
}

// The digits of r = a + b p as a ladder leaves them: |a| <= p/2 (1 + 2^-20) after a window
// step (0 <= a < p if none ran) and |b| <= 1.1 p.  Then r = 1 (mod p), the FLT condition,
// holds exactly when a == 1, and q = (r - 1) / p = b mod p takes at most two corrections.
// q < 2^51 is read from the bits of q + 2^52 (exponent fixed, the mantissa is q).  Returns
// false, leaving q alone, when a != 1.
__device__ __forceinline__ bool f64_quotient(double a, double b, double pd, u64& q) {
  if (a != 1.0) return false;
  double t = b < 0.0 ? b + pd : b;                      // (-0.1 p, 1.1 p]
  t = t < 0.0 ? t + pd : (t >= pd ? t - pd : t);
  const double s = t + 4503599627370496.0;
  u64 bits;
  __builtin_memcpy(&bits, &s, sizeof bits);
  q = bits & 0xFFFFFFFFFFFFFull;
  return true;
}

// x *= c for an integer 0 < c < 2^32 (any base's window multiplier b^d): a c and b c are
// split exactly into two doubles (h + l, FMA), each reduced by one rounded quotient, and the
// b digit takes the carry k1 and one more reduction.  Afterwards |a|, |b| <= p/2 (+ 2^-21 p).
__device__ __forceinline__ void f64_mul_small(double& a, double& b, double c, const F64Mod& m) {
  const double h1 = a * c;
  const double l1 = __builtin_fma(a, c, -h1);
  const double k1 = round_mul(h1, m.ip);                    // |k1| < 2^32
  const double a1 = __builtin_fma(-k1, m.p, h1) + l1;       // a c - k1 p
  const double h2 = b * c;
  const double l2 = __builtin_fma(b, c, -h2);
  const double k2 = round_mul(h2, m.ip);
  const double b1 = (__builtin_fma(-k2, m.p, h2) + l2) + k1;   // b c - k2 p + k1
  const double k3 = round_mul(b1, m.ip);
  b = __builtin_fma(-k3, m.p, b1);
  a = a1;
}

// The same for p >= 2^32, one reduction shorter: |k1| <= 0.59 c + 1 < 0.6 p, so
// b' = (b c - k2 p + l2) + k1 stays within 1.1 p, inside the squaring's bound |b| <= 1.9 p.
__device__ __forceinline__ void f64_mul_small_big(double& a, double& b, double c, const F64Mod& m) {
  const double h1 = a * c;
  const double l1 = __builtin_fma(a, c, -h1);
  const double k1 = round_mul(h1, m.ip);
  const double a1 = __builtin_fma(-k1, m.p, h1) + l1;
  const double h2 = b * c;
  const double l2 = __builtin_fma(b, c, -h2);
  const double k2 = round_mul(h2, m.ip);
  b = (__builtin_fma(-k2, m.p, h2) + l2) + k1;
  a = a1;
}

// Two ladders with window W, for the base of pw0 on (p0, e0) and the base of pw1 on (p1, e1):
// pw[d] = b^d < 2^32 for d < 2^W (one table twice for one base on two primes, or two tables
// on one prime).  x = b^(leading digit), then per window W squarings and one multiplication
// by b^d (BIG: both p >= 2^32).  The digits are left raw, as fp64_pow2x2 leaves them.
template <int W, bool BIG = false>
__device__ __forceinline__ void fp64_powc_x2_r(double pd0, double ip0, u64 e0, double pd1, double ip1, u64 e1,
                                               int ebits, const u32* pw0, const u32* pw1, double& a0, double& b0,
                                               double& a1, double& b1) {
  F64Mod m0, m1;
  m0.p = pd0;
  m0.ip = ip0;
  m1.p = pd1;
  m1.ip = ip1;
  const int nw = (ebits + W - 1) / W;
  int sh = (nw - 1) * W;
  const u32 c0 = pw0[(int)(e0 >> sh) & ((1 << W) - 1)], c1 = pw1[(int)(e1 >> sh) & ((1 << W) - 1)];
  a0 = 1.0; b0 = 0.0; a1 = 1.0; b1 = 0.0;
  if ((double)c0 < pd0) a0 = (double)c0;
  else f64_mul_small(a0, b0, (double)c0, m0);
  if ((double)c1 < pd1) a1 = (double)c1;
  else f64_mul_small(a1, b1, (double)c1, m1);
  for (sh -= W; sh >= 0; sh -= W) {
#pragma unroll
    for (int s = 0; s < W; s++) {
      f64_sqr(a0, b0, m0);
      f64_sqr(a1, b1, m1);
    }
    if (BIG) {
      f64_mul_small_big(a0, b0, (double)pw0[(int)((e0 >> sh) & ((1u << W) - 1))], m0);
      f64_mul_small_big(a1, b1, (double)pw1[(int)((e1 >> sh) & ((1u << W) - 1))], m1);
    } else {
      f64_mul_small(a0, b0, (double)pw0[(int)((e0 >> sh) & ((1u << W) - 1))], m0);
      f64_mul_small(a1, b1, (double)pw1[(int)((e1 >> sh) & ((1u << W) - 1))], m1);
    }
  }
}

template <int W, bool BIG = false>
__device__ __forceinline__ void fp64_powc_x2(double pd0, u64 e0, double pd1, u64 e1, int ebits, const u32* pw0,
                                             const u32* pw1, double& a0, double& b0, double& a1, double& b1) {
  fp64_powc_x2_r<W, BIG>(pd0, 1.0 / pd0, e0, pd1, 1.0 / pd1, e1, ebits, pw0, pw1, a0, b0, a1, b1);
}

// ---------------------------------------------------------------------------------
// 32-bit limb Montgomery path
// ---------------------------------------------------------------------------------

// REDC of a double-width t < N R: x = t / R mod N, one low word cleared per row.  The
// result is below 2N; FULL subtracts N once (keeping the carry bit past 2^(32L)).
template <int L, bool FULL>
__device__ __forceinline__ void mont_redc(u32 (&t)[2 * L], u32 (&x)[L], const u32 (&n)[L], u32 ninv) {
  u32 hi = 0;   // the carry pending into word i + L + 1
#pragma unroll
  for (int i = 0; i < L; i++) {
    const u32 mq = t[i] * ninv;
    u64 c = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
      c = (u64)mq * n[j] + t[i + j] + (c >> 32);
      t[i + j] = (u32)c;
    }
    const u64 s = (u64)t[i + L] + (c >> 32) + hi;
    t[i + L] = (u32)s;
    hi = (u32)(s >> 32);
  }
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = t[L + i];
  if (FULL) {
    // result = hi 2^(32L) + x < 2N: subtract N once if it is >= N
    u32 d[L];
    u64 br = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
      const u64 s = (u64)x[i] - n[i] - br;
      d[i] = (u32)s;
      br = (s >> 32) & 1;
    }
    const bool ge = hi || !br;
#pragma unroll
    for (int i = 0; i < L; i++) x[i] = ge ? d[i] : x[i];
  }
}

// Montgomery squaring t = x^2 / R mod N (R = 2^(32L)), ninv = -N^-1 mod 2^32.
// Lazy (FULL = false): x < 2N, N < R/4 gives a result < 2N and no final subtraction.
// FULL: x < N < R; the result (< 2N < 2^(32L+1)) keeps its carry bit and is reduced.
template <int L, bool FULL>
__device__ __forceinline__ void mont_sqr(u32 (&x)[L], const u32 (&n)[L], u32 ninv) {
  u32 t[2 * L];
#pragma unroll
  for (int k = 0; k < 2 * L; k++) t[k] = 0;
  // cross products x_i x_j, i < j
#pragma unroll
  for (int i = 0; i < L - 1; i++) {
    u64 c = 0;
#pragma unroll
    for (int j = i + 1; j < L; j++) {
      c = (u64)x[i] * x[j] + t[i + j] + (c >> 32);
      t[i + j] = (u32)c;
    }
    t[i + L] = (u32)(c >> 32);
  }
  // double them
#pragma unroll
  for (int k = 2 * L - 1; k > 0; k--) t[k] = (t[k] << 1) | (t[k - 1] >> 31);
  t[0] <<= 1;
  // add the squares x_i^2 at word 2i
  {
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
      const u64 sq = (u64)x[i] * x[i];
      c = (u64)t[2 * i] + (u32)sq + (c >> 32);
      t[2 * i] = (u32)c;
      c = (u64)t[2 * i + 1] + (u32)(sq >> 32) + (c >> 32);
      t[2 * i + 1] = (u32)c;
    }
  }
  mont_redc<L, FULL>(t, x, n, ninv);
}

// x = x y / R mod N for x < 2N (lazy) or x < N (FULL) and y < N: the same bounds as mont_sqr.
template <int L, bool FULL>
__device__ __forceinline__ void mont_mul(u32 (&x)[L], const u32 (&y)[L], const u32 (&n)[L], u32 ninv) {
  u32 t[2 * L];
#pragma unroll
  for (int k = 0; k < 2 * L; k++) t[k] = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    u64 c = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
      c = (u64)x[i] * y[j] + t[i + j] + (c >> 32);
      t[i + j] = (u32)c;
    }
    t[i + L] = (u32)(c >> 32);   // row i is the first to write word i + L
  }
  mont_redc<L, FULL>(t, x, n, ninv);
}

// x = 2x mod M (M = 2N in the lazy domain [0, 2N), M = N for FULL with carry-out).
template <int L>
__device__ __forceinline__ void mont_dbl(u32 (&x)[L], const u32 (&M)[L]) {
  u32 y[L];
  const u32 top = x[L - 1] >> 31;
#pragma unroll
  for (int k = L - 1; k > 0; k--) y[k] = (x[k] << 1) | (x[k - 1] >> 31);
  y[0] = x[0] << 1;
  u32 d[L];
  u64 br = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const u64 s = (u64)y[i] - M[i] - br;
    d[i] = (u32)s;
    br = (s >> 32) & 1;
  }
  const bool ge = top || !br;
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = ge ? d[i] : y[i];
}

// REDC of a single-width value x < 2N: x / R mod N, fully reduced.
template <int L>
__device__ __forceinline__ void mont_out(u32 (&x)[L], const u32 (&n)[L], u32 ninv) {
  u32 t[2 * L];
#pragma unroll
  for (int k = 0; k < L; k++) { t[k] = x[k]; t[L + k] = 0; }
  u32 hi = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    const u32 mq = t[i] * ninv;
    u64 c = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
      c = (u64)mq * n[j] + t[i + j] + (c >> 32);
      t[i + j] = (u32)c;
    }
    const u64 s = (u64)t[i + L] + (c >> 32) + hi;
    t[i + L] = (u32)s;
    hi = (u32)(s >> 32);
  }
  u32 d[L];
  u64 br = 0;
#pragma unroll
  for (int i = 0; i < L; i++) {
    x[i] = t[L + i];
    const u64 s = (u64)x[i] - n[i] - br;
    d[i] = (u32)s;
    br = (s >> 32) & 1;
  }
  const bool ge = hi || !br;
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = ge ? d[i] : x[i];
}

// r = 2^e mod N through the L-limb Montgomery ladder; r has L limbs (N = p^2).
// ebits = the wave's longest exponent length (shorter exponents see leading zeros).
template <int L, bool FULL>
__device__ __forceinline__ void mont_pow2(u64 p, u64 e, int ebits, u32 (&r)[L]) {
  // N = p^2 as L limbs
  u32 n[L];
  {
    const u64 lo = p * p;
    const u64 hi = __umul64hi(p, p);
    const u32 w4[4] = {(u32)lo, (u32)(lo >> 32), (u32)hi, (u32)(hi >> 32)};
#pragma unroll
    for (int i = 0; i < L; i++) n[i] = i < 4 ? w4[i] : 0u;
  }
  const u64 pinv = inv64(p);
  const u32 ninv = 0u - (u32)pinv * (u32)pinv;
  // M = modulus of the doubling: 2N (lazy) or N (full)
  u32 M[L];
  {
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
      M[i] = FULL ? n[i] : ((n[i] << 1) | c);
      c = n[i] >> 31;
    }
  }
  // x = 2^(bits(N) - 1) < N, then doublings mod N up to 2^(32L + 1) = Montgomery form of 2
  int topw = L - 1;
  while (topw > 0 && n[topw] == 0) topw--;
  const int nb = 32 * topw + (32 - __builtin_clz(n[topw]));  // bits(N)
  u32 x[L];
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = 0;
  const int hb = nb - 1;
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = (i == (hb >> 5)) ? (1u << (hb & 31)) : 0u;
  for (int k = hb; k < 32 * L + 1; k++) mont_dbl<L>(x, n);   // reduce mod N while building
  // ladder over the exponent bits below the top one
  for (int i = ebits - 2; i >= 0; i--) {
    const bool lead = i >= 63 - (int)__builtin_clzll(e);    // above this lane's top bit
    if (!lead) {
      mont_sqr<L, FULL>(x, n, ninv);
      if ((e >> i) & 1) mont_dbl<L>(x, M);
    }
  }
  mont_out<L>(x, n, ninv);
#pragma unroll
  for (int i = 0; i < L; i++) r[i] = x[i];
}

// r = b^e mod N for any base b (N = p^2) by the plain binary ladder: one squaring per bit and
// one Montgomery product by B = b R mod N per set bit.  This is a second arithmetic for the
// bases other than 2, independent of the window ladders of the FP64 and radix-2^28 paths
// (no power table, no window, 32-bit limbs): the double check recomputes with it.
template <int L, bool FULL>
__device__ __forceinline__ void mont_powb(u64 p, u64 e, int ebits, u32 b, u32 (&r)[L]) {
  const u64 lo = p * p;
  const u64 hi = __umul64hi(p, p);
  u32 n[L];
  {
    const u32 w4[4] = {(u32)lo, (u32)(lo >> 32), (u32)hi, (u32)(hi >> 32)};
#pragma unroll
    for (int i = 0; i < L; i++) n[i] = i < 4 ? w4[i] : 0u;
  }
  const u64 pinv = inv64(p);
  const u32 ninv = 0u - (u32)pinv * (u32)pinv;
  // B = (b mod N) doubled 32L times mod N: every step keeps it below N
  const u64 bm = hi || b < lo ? (u64)b : (u64)b % lo;
  u32 B[L];
#pragma unroll
  for (int i = 0; i < L; i++) B[i] = i == 0 ? (u32)bm : i == 1 ? (u32)(bm >> 32) : 0u;
  for (int k = 0; k < 32 * L; k++) mont_dbl<L>(B, n);
  u32 x[L];
#pragma unroll
  for (int i = 0; i < L; i++) x[i] = B[i];
  for (int i = ebits - 2; i >= 0; i--) {
    const bool lead = i >= 63 - (int)__builtin_clzll(e);    // above this lane's top bit
    if (!lead) {
      mont_sqr<L, FULL>(x, n, ninv);
      if ((e >> i) & 1) mont_mul<L, FULL>(x, B, n, ninv);
    }
  }
  mont_out<L>(x, n, ninv);
#pragma unroll
  for (int i = 0; i < L; i++) r[i] = x[i];
}

// ---------------------------------------------------------------------------------
// radix-2^28 Montgomery, 5 limbs, R = 2^140 (any odd p with 2^47 <= p < 2^64)
// ---------------------------------------------------------------------------------
// Limbs of 28 bits in 32-bit registers; every partial product (< 2^59) is one
// v_mad_u64_u32 into a 64-bit column sum, and carries move once per column, so no
// product needs its own carry chain.  N = p^2 < 2^128 leaves 12 spare bits below R.
// The squaring's REDC multipliers keep all 32 bits of m_j = T_j ninv (only their low 28
// bits matter for clearing a column), so m = sum m_j 2^(28 j) < 2^144 (1 + 2^-27) and any
// input x < 64 N squares to (x^2 + m N) / R < N (4096 N / R + 16.001) < 17.01 N.  The base-2
// ladder therefore doubles without reducing (below 35 N; limbs 0..3 then hold 29 bits and
// limb 4 22 bits, so the squaring's products stay < 2^59, the REDC products < 2^60 and every
// column sum < 2^63), and only the final REDC, with 28-bit multipliers, reduces fully.  This
// also covers p >= 2^63, where N reaches 2^128 and a 32-bit-limb REDC needs its 129th bit.

#define M28 0x0FFFFFFFu

// One REDC row: clear column T0 (mod 2^28) with m N, carry T0 into T1.  M28_ROW takes m
// below 2^28 (the final, full reduction); M28_ROW_LAZY keeps m's top four bits.
#define M28_ROW(T0, T1, T2, T3, T4)                  \
  m = ((u32)T0 * ninv) & M28;                      \
  T0 += (u64)m * n[0];                             \
  T1 += (u64)m * n[1] + (T0 >> 28);                \
  T2 += (u64)m * n[2];                             \
  T3 += (u64)m * n[3];                             \
  T4 += (u64)m * n[4];
#define M28_ROW_LAZY(T0, T1, T2, T3, T4)             \
  m = (u32)T0 * ninv;                              \
  T0 += (u64)m * n[0];                             \
  T1 += (u64)m * n[1] + (T0 >> 28);                \
  T2 += (u64)m * n[2];                             \
  T3 += (u64)m * n[3];                             \
  T4 += (u64)m * n[4];

// x = x^2 / R mod N (lazy: input < 64 N, result < 17.01 N with limbs 0..3 below 2^28)
__device__ __forceinline__ void m28_sqr(u32 (&x)[5], const u32 (&n)[5], u32 ninv) {
  const u32 d1 = x[1] << 1, d2 = x[2] << 1, d3 = x[3] << 1, d4 = x[4] << 1;
  u64 t0 = (u64)x[0] * x[0];
  u64 t1 = (u64)x[0] * d1;
  u64 t2 = (u64)x[0] * d2 + (u64)x[1] * x[1];
  u64 t3 = (u64)x[0] * d3 + (u64)x[1] * d2;
  u64 t4 = (u64)x[0] * d4 + (u64)x[1] * d3 + (u64)x[2] * x[2];
  u64 t5 = (u64)x[1] * d4 + (u64)x[2] * d3;
  u64 t6 = (u64)x[2] * d4 + (u64)x[3] * x[3];
  u64 t7 = (u64)x[3] * d4;
  u64 t8 = (u64)x[4] * x[4];
  u32 m;
  M28_ROW_LAZY(t0, t1, t2, t3, t4)
  M28_ROW_LAZY(t1, t2, t3, t4, t5)
  M28_ROW_LAZY(t2, t3, t4, t5, t6)
  M28_ROW_LAZY(t3, t4, t5, t6, t7)
  M28_ROW_LAZY(t4, t5, t6, t7, t8)
  x[0] = (u32)t5 & M28;
  t6 += t5 >> 28;
  x[1] = (u32)t6 & M28;
  t7 += t6 >> 28;
  x[2] = (u32)t7 & M28;
  t8 += t7 >> 28;
  x[3] = (u32)t8 & M28;
  x[4] = (u32)(t8 >> 28);
}

// REDC of x < 64 N: x / R mod N, fully reduced, as four 32-bit words.
__device__ __forceinline__ void m28_out(const u32 (&x)[5], const u32 (&n)[5], u32 ninv, u32 (&r)[4]) {
  u64 t0 = x[0], t1 = x[1], t2 = x[2], t3 = x[3], t4 = x[4], t5 = 0, t6 = 0, t7 = 0, t8 = 0;
  u32 m;
  M28_ROW(t0, t1, t2, t3, t4)
  M28_ROW(t1, t2, t3, t4, t5)
  M28_ROW(t2, t3, t4, t5, t6)
  M28_ROW(t3, t4, t5, t6, t7)
  M28_ROW(t4, t5, t6, t7, t8)
  t6 += t5 >> 28;
  t7 += t6 >> 28;
  t8 += t7 >> 28;
  // y = (x + m N) / R <= N, in limbs t5..t8 (t8 < 2^44)
  const unsigned __int128 y = ((unsigned __int128)t8 << 84) | ((unsigned __int128)((u32)t7 & M28) << 56) |
                              ((unsigned __int128)((u32)t6 & M28) << 28) | ((u32)t5 & M28);
  const unsigned __int128 N = ((unsigned __int128)n[4] << 112) | ((unsigned __int128)n[3] << 84) |
                              ((unsigned __int128)n[2] << 56) | ((unsigned __int128)n[1] << 28) | n[0];
  const unsigned __int128 z = y >= N ? y - N : y;
  r[0] = (u32)z;
  r[1] = (u32)(z >> 32);
  r[2] = (u32)(z >> 64);
  r[3] = (u32)(z >> 96);
}
#undef M28_ROW
#undef M28_ROW_LAZY

// N = p^2 as five 28-bit limbs, -N^-1 mod 2^28, R mod N (R = 2^140) and 1/N in FP64.
struct M28Mod {
  u32 n[5];
  u32 ninv;
  unsigned __int128 rmodn;
  double invN;
};

__device__ __forceinline__ void m28_setup(u64 p, M28Mod& m) {
  typedef unsigned __int128 u128;
  const u128 N = (u128)p * p;
  for (int i = 0; i < 5; i++) m.n[i] = (u32)(N >> (28 * i)) & M28;
  const u64 pinv = inv64(p);
  m.ninv = (0u - (u32)pinv * (u32)pinv) & M28;
  m.invN = 1.0 / ((double)p * (double)p);
  // R mod N = 2^140 - Q N with Q = floor(2^140 / N) < 2^46 estimated in FP64 (relative
  // error < 2^-50, so off by at most one); the 192-bit difference is then brought into
  // [0, N) by at most two additions or subtractions of N.
  const u64 Q = (u64)(0x1p140 * m.invN);
  const u128 A = (u128)Q * (u64)N;                   // Q * N_lo
  const u128 B = (u128)Q * (u64)(N >> 64);           // Q * N_hi
  const u128 low = A + (B << 64);                    // Q N mod 2^128
  const long long top = (long long)((u64)(B >> 64) + (low < A ? 1u : 0u));   // Q N >> 128
  u128 vlow = (u128)0 - low;                         // 2^140 - Q N = vtop 2^128 + vlow
  long long vtop = 4096 - top - (low != 0 ? 1 : 0);
  while (vtop < 0) {
    const u128 t = vlow + N;
    vtop += t < vlow ? 1 : 0;
    vlow = t;
  }
  while (vtop > 0 || vlow >= N) {
    vtop -= vlow < N ? 1 : 0;
    vlow -= N;
  }
  m.rmodn = vlow;
}

// x = x c mod N for an integer c < 2^32 (x < 64 N on entry, < 3 N on exit).  The
// Montgomery form of a product with a plain integer is that form times the integer, so no
// REDC is needed: T = x c < 2^38 N, its quotient by N is estimated in FP64 from the top
// three limbs (the lower ones move T / N by < 2^-16; the estimate is off by at most one)
// and one less than the estimate is subtracted, column by column with signed carries.
__device__ __forceinline__ void m28_mul_small(u32 (&x)[5], u32 c, const M28Mod& m) {
  u64 t0 = (u64)x[0] * c, t1 = (u64)x[1] * c, t2 = (u64)x[2] * c, t3 = (u64)x[3] * c, t4 = (u64)x[4] * c;
  t1 += t0 >> 28;
  t2 += t1 >> 28;
  t3 += t2 >> 28;
  t4 += t3 >> 28;
  const u64 l5 = t4 >> 28;                               // < 2^39
  const u32 l3 = (u32)t3 & M28, l4 = (u32)t4 & M28;
  const double td = ((double)l5 * 0x1p140 + (double)l4 * 0x1p112) + (double)l3 * 0x1p84;
  double qd = __builtin_floor(td * m.invN);
  qd = qd >= 1.0 ? qd - 1.0 : 0.0;
  const u64 q = (u64)qd;                                  // < 2^38
  const u64 q0 = q & M28, q1 = q >> 28;
  const long long c0 = (long long)((u32)t0 & M28) - (long long)(q0 * m.n[0]);
  const long long c1 = (long long)((u32)t1 & M28) - (long long)(q0 * m.n[1] + q1 * m.n[0]) + (c0 >> 28);
  const long long c2 = (long long)((u32)t2 & M28) - (long long)(q0 * m.n[2] + q1 * m.n[1]) + (c1 >> 28);
  const long long c3 = (long long)l3 - (long long)(q0 * m.n[3] + q1 * m.n[2]) + (c2 >> 28);
  const long long c4 = (long long)l4 - (long long)(q0 * m.n[4] + q1 * m.n[3]) + (c3 >> 28);
  // the result is below 3N < 2^130: limb 4 takes bits 112..139 and nothing is left above
  x[0] = (u32)c0 & M28;
  x[1] = (u32)c1 & M28;
  x[2] = (u32)c2 & M28;
  x[3] = (u32)c3 & M28;
  x[4] = (u32)c4 & M28;
}

// r = 2^e mod N = p^2 as four 32-bit words, for odd p >= 2^47.  ebits = the wave's longest
// exponent length (shorter exponents skip their leading positions).  The exponent's five
// leading bits t (16 <= t < 32) enter at once: the Montgomery form of 2^t is (R mod N) 2^t,
// one m28_mul_small in place of four squarings.  Each later bit is a squaring and, for a
// one, a doubling by limb shifts (below 35 N, inside the squaring's 64 N).
__device__ __forceinline__ void m28_pow2(u64 p, u64 e, int ebits, u32 (&r)[4]) {
  M28Mod m;
  m28_setup(p, m);
  u32 x[5];                                          // Montgomery form of 1: R mod N < N
  for (int i = 0; i < 5; i++) x[i] = (u32)(m.rmodn >> (28 * i)) & M28;
  const int tb = 63 - __builtin_clzll(e);            // >= 46 for p >= 2^47
  m28_mul_small(x, 1u << (u32)(e >> (tb - 4)), m);   // 2^t R mod N, < 3N
  for (int i = ebits - 6; i >= 0; i--) {
    if (i > tb - 5) continue;
    m28_sqr(x, m.n, m.ninv);
    const u32 bit = (u32)(e >> i) & 1u;
    x[0] <<= bit; x[1] <<= bit; x[2] <<= bit; x[3] <<= bit; x[4] <<= bit;
  }
  m28_out(x, m.n, m.ninv, r);
}

// r = b^e mod N = p^2 (four 32-bit words) for a base with window W: pw[d] = b^d < 2^32.
template <int W>
__device__ __forceinline__ void m28_powc(u64 p, u64 e, int ebits, const u32* pw, u32 (&r)[4]) {
  M28Mod m;
  m28_setup(p, m);
  u32 x[5];                                          // Montgomery form of 1: R mod N < N
  for (int i = 0; i < 5; i++) x[i] = (u32)(m.rmodn >> (28 * i)) & M28;
  const int nw = (ebits + W - 1) / W;
  int sh = (nw - 1) * W;
  // e < 2^ebits for every active lane, so the leading digit is below 2^W; the mask keeps an
  // idle lane's longer exponent inside the table as well
  m28_mul_small(x, pw[(int)(e >> sh) & ((1 << W) - 1)], m);
  for (sh -= W; sh >= 0; sh -= W) {
#pragma unroll
    for (int s = 0; s < W; s++) m28_sqr(x, m.n, m.ninv);
    m28_mul_small(x, pw[(int)((e >> sh) & ((1u << W) - 1))], m);
  }
  m28_out(x, m.n, m.ninv, r);
}

// ---------------------------------------------------------------------------------
// base-2 strong probable-prime test of an odd n >= 3, n < 2^64
// ---------------------------------------------------------------------------------
// Montgomery arithmetic mod n with R = 2^64 in two 32-bit limbs, interleaving the rows of
// a b with the REDC rows (each step one v_mad_u64_u32: 32 x 32 + 64 bits): 8 multiply-adds
// and 2 low multiplies per product.  n may exceed 2^63, so the result keeps its carry bit
// and is reduced to [0, n).

// a b / 2^64 mod n for a, b < n, ninv = -n^-1 mod 2^64 (its low word is -n^-1 mod 2^32)
__device__ __forceinline__ u64 mont64_mul(u64 a, u64 b, u64 n, u64 ninv) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  const u32 n0 = (u32)n, n1 = (u32)(n >> 32), ni = (u32)ninv;
  // row 0: t = a0 b
  u64 c = (u64)a0 * b0;
  u32 t0 = (u32)c;
  c = (u64)a0 * b1 + (c >> 32);
  u32 t1 = (u32)c, t2 = (u32)(c >> 32);
  // clear word 0: t = (t + m n) / 2^32
  u32 m = t0 * ni;
  c = (u64)m * n0 + t0;
  c = (u64)m * n1 + t1 + (c >> 32);
  t0 = (u32)c;
  c = (u64)t2 + (c >> 32);
  t1 = (u32)c;
  t2 = (u32)(c >> 32);
  // row 1: t += a1 b
  c = (u64)a1 * b0 + t0;
  t0 = (u32)c;
  c = (u64)a1 * b1 + t1 + (c >> 32);
  t1 = (u32)c;
  const u64 top = (u64)t2 + (c >> 32);
  // clear word 0 again; r = (r1 2^32 + r0) < 2n < 2^65
  m = t0 * ni;
  c = (u64)m * n0 + t0;
  c = (u64)m * n1 + t1 + (c >> 32);
  const u64 r1 = top + (c >> 32);
  const u64 r = (r1 << 32) | (u32)c;
  return ((r1 >> 32) || r >= n) ? r - n : r;
}

// a^2 / 2^64 mod n: mont64_mul(a, a) with the cross product a0 a1 formed once (7
// multiply-adds instead of 8)
__device__ __forceinline__ u64 mont64_sqr(u64 a, u64 n, u64 ninv) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32);
  const u32 n0 = (u32)n, n1 = (u32)(n >> 32), ni = (u32)ninv;
  const u64 x = (u64)a0 * a1;
  // row 0: t = a0 a
  u64 c = (u64)a0 * a0;
  u32 t0 = (u32)c;
  c = x + (c >> 32);
  u32 t1 = (u32)c, t2 = (u32)(c >> 32);
  u32 m = t0 * ni;
  c = (u64)m * n0 + t0;
  c = (u64)m * n1 + t1 + (c >> 32);
  t0 = (u32)c;
  c = (u64)t2 + (c >> 32);
  t1 = (u32)c;
  t2 = (u32)(c >> 32);
  // row 1: t += a1 a = a1 a0 + a1 a1 2^32
  c = x + t0;
  t0 = (u32)c;
  c = (u64)a1 * a1 + t1 + (c >> 32);
  t1 = (u32)c;
  const u64 top = (u64)t2 + (c >> 32);
  m = t0 * ni;
  c = (u64)m * n0 + t0;
  c = (u64)m * n1 + t1 + (c >> 32);
  const u64 r1 = top + (c >> 32);
  const u64 r = (r1 << 32) | (u32)c;
  return ((r1 >> 32) || r >= n) ? r - n : r;
}

__device__ __forceinline__ u64 mont64_dbl(u64 x, u64 n) {
  const bool c = x >> 63;
  const u64 y = x << 1;
  return (c || y >= n) ? y - n : y;
}

// m^-1 mod 2^32 for odd m (Newton's iteration, exact after five steps)
constexpr u32 inv32c(u32 m) {
  u32 x = m;
  for (int k = 0; k < 5; k++) x *= 2u - m * x;
  return x;
}

// m | x for an odd constant m and any x < 2^32: x m^-1 mod 2^32 equals x / m exactly when m
// divides x, and every other value exceeds (2^32 - 1) / m.  One multiply and one compare.
template <u32 M>
__device__ __forceinline__ bool divides32(u32 x) {
  constexpr u32 I = inv32c(M);
  return x * I <= 0xFFFFFFFFu / M;
}

// A necessary condition for an odd composite n to be a base-2 strong probable prime.
// Write n - 1 = d 2^s.  Either 2^d = 1 (mod n), and then ord_r(2) | d for every prime factor
// r of n, or 2^(d 2^t) = -1 (mod n) for some t < s, and then ord_r(2) | d 2^(t+1) but not
// d 2^t.  Either way ord_r(2) | n - 1 and all prime factors share the 2-adic valuation of
// ord_r(2).  Checked for the factors r < 100; false means n is certainly not a strong
// probable prime.
// Branch-free: the residues of n modulo five products of the small primes (each below
// 2^32) give, by divides32, both the prime factors r < 100 of n and the congruences n = 1
// mod the odd prime powers in ord_r(2); those modulo 4, 8 and 16 are n's low bits.  The
// 2-adic valuations of the orders of the factors found are collected as bits of vm, which
// must hold at most one bit.  Every lane runs the same instructions, so a wavefront never
// splits here.  psp_possible_res takes n's residues (gi = n mod PSP_Mi, g5 = n mod 97) and
// its low four bits n4, which a caller walking consecutive numbers can advance cheaply.
constexpr u32 PSP_M1 = 9u * 5 * 7 * 11 * 13 * 17 * 19 * 23;   // 334639305
constexpr u32 PSP_M2 = 29u * 31 * 37 * 41 * 43;               // 58642669
constexpr u32 PSP_M3 = 47u * 53 * 59 * 61 * 67;               // 600662303
constexpr u32 PSP_M4 = 71u * 73 * 79 * 83 * 89;               // 3024658859
__device__ __forceinline__ bool psp_possible_res(u32 n4, u32 g1, u32 g2, u32 g3, u32 g4, u32 g5) {
  constexpr u32 M1 = PSP_M1, M2 = PSP_M2;
  // n = 1 modulo 4, 8, 16 and modulo odd divisors of M1 and M2 (g + M - 1 < 2^32 avoids the
  // wrap of g - 1 at g = 0)
  const bool e4 = (n4 & 3u) == 1u, e8 = (n4 & 7u) == 1u, e16 = n4 == 1u;
  const u32 h1 = g1 + (M1 - 1u), h2 = g2 + (M2 - 1u);
  const bool e3 = divides32<3>(h1), e9 = divides32<9>(h1), e5 = divides32<5>(h1), e7 = divides32<7>(h1),
             e11 = divides32<11>(h1), e13 = divides32<13>(h1), e23 = divides32<23>(h1), e29 = divides32<29>(h2),
             e41 = divides32<41>(h2);
  bool ok = true;
  u32 vm = 0;
  // r | n requires n = 1 (mod ord_r(2)); v = the 2-adic valuation of ord_r(2)
  auto factor = [&](bool d, bool cond, int v) {
    ok = ok && (!d || cond);
    vm |= (d ? 1u : 0u) << v;
  };
  factor(divides32<3>(g1), true, 1);                 // ord 2
  factor(divides32<5>(g1), e4, 2);                   // 4
  factor(divides32<7>(g1), e3, 0);                   // 3
  factor(divides32<11>(g1), e5, 1);                  // 10
  factor(divides32<13>(g1), e4 && e3, 2);            // 12
  factor(divides32<17>(g1), e8, 3);                  // 8
  factor(divides32<19>(g1), e9, 1);                  // 18
  factor(divides32<23>(g1), e11, 0);                 // 11
  factor(divides32<29>(g2), e4 && e7, 2);            // 28
  factor(divides32<31>(g2), e5, 0);                  // 5
  factor(divides32<37>(g2), e4 && e9, 2);            // 36
  factor(divides32<41>(g2), e4 && e5, 2);            // 20
  factor(divides32<43>(g2), e7, 1);                  // 14
  factor(divides32<47>(g3), e23, 0);                 // 23
  factor(divides32<53>(g3), e4 && e13, 2);           // 52
  factor(divides32<59>(g3), e29, 1);                 // 58
  factor(divides32<61>(g3), e4 && e3 && e5, 2);      // 60
  factor(divides32<67>(g3), e3 && e11, 1);           // 66
  factor(divides32<71>(g4), e5 && e7, 0);            // 35
  factor(divides32<73>(g4), e9, 0);                  // 9
  factor(divides32<79>(g4), e3 && e13, 0);           // 39
  factor(divides32<83>(g4), e41, 1);                 // 82
  factor(divides32<89>(g4), e11, 0);                 // 11
  factor(g5 == 0u, e16 && e3, 4);                    // 48
  return ok && (vm & (vm - 1u)) == 0u;
}

__device__ __forceinline__ bool psp_possible(u64 n) {
  return psp_possible_res((u32)n & 15u, (u32)(n % PSP_M1), (u32)(n % PSP_M2), (u32)(n % PSP_M3), (u32)(n % PSP_M4),
                          (u32)(n % 97u));
}

// x 2^k mod n for x < n and 0 <= k < 32 (a Montgomery form times 2^k is the form of the
// product).  The quotient (below 2^31) comes from an FP64 estimate with invn = 1/n rounded
// (relative error below 2^-51, so within one of the truth); one less is subtracted in
// 128-bit integers, leaving [0, 3n), then at most two subtractions of n.
__device__ __forceinline__ u64 mul_pow2_mod(u64 x, int k, u64 n, double invn) {
  typedef unsigned __int128 u128;
  double qd = __builtin_floor(__builtin_ldexp((double)x, k) * invn);
  qd = qd >= 1.0 ? qd - 1.0 : 0.0;
  u128 r = ((u128)x << k) - (u128)(u64)qd * n;
  r = r >= n ? r - n : r;
  r = r >= n ? r - n : r;
  return (u64)r;
}

// true iff n is a base-2 strong probable prime (n odd, n >= 3)
__device__ __forceinline__ bool sprp2_64(u64 n) {
  const u64 ninv = 0 - inv64(n);
  const double invn = 1.0 / (double)n;
  const int s = __builtin_ctzll(n - 1);
  const u64 d = (n - 1) >> s;
  // one = R mod n: 2^(bits(n) - 1) doubled up to 2^64
  const int nb = 64 - __builtin_clzll(n);
  u64 one = 1ull << (nb - 1);
  for (int k = nb - 1; k < 64; k++) one = mont64_dbl(one, n);
  const u64 mone = n - one;                       // Montgomery form of n - 1
  // x = 2^d in Montgomery form, left to right in 5-bit digits of d (aligned to its bit 0):
  // the top digit t gives (R mod n) 2^t, each further digit five squarings and one
  // multiplication by 2^digit
  const int db = 64 - __builtin_clzll(d);
  int sh = (db - 1) / 5 * 5;
  u64 x = mul_pow2_mod(one, (int)(d >> sh), n, invn);
  for (sh -= 5; sh >= 0; sh -= 5) {
#pragma unroll
    for (int k = 0; k < 5; k++) x = mont64_sqr(x, n, ninv);
    x = mul_pow2_mod(x, (int)((d >> sh) & 31u), n, invn);
  }
  if (x == one || x == mone) return true;
  for (int r = 1; r < s; r++) {
    x = mont64_sqr(x, n, ninv);
    if (x == mone) return true;
  }
  return false;
}

// ---------------------------------------------------------------------------------
// Strong Lucas test (Selfridge's parameters) and BPSW, odd n < 2^64
// ---------------------------------------------------------------------------------
//
// Baillie-Wagstaff (Math. Comp. 35, 1980): P = 1, Q = (1 - D)/4 with D the first of
// 5, -7, 9, -11, ... whose Jacobi symbol (D/n) is -1.  With n + 1 = d 2^s, n is a strong
// Lucas probable prime when U_d = 0 or V_(d 2^r) = 0 (mod n) for some 0 <= r < s.  BPSW is
// this test after the base-2 strong test; no composite below 2^64 passes both (checked
// against the complete list of base-2 Fermat pseudoprimes below 2^64, Feitsma and
// Gilchrist), so below 2^64 it decides primality.  The reference runs the same pair in
// A350826/cuda/a350826_cuda.cu:768-794 (base 2) and :855-898 (Lucas).
//
// Here only V is carried: the ladder keeps V_k, V_(k+1) and Q^k (the same four products
// for a 0 and a 1 bit, so the lanes of a wavefront never diverge), and U_d = 0 is read
// from D U_d = 2 V_(d+1) - V_d (gcd(D, n) = 1 since (D/n) = -1).

__device__ __forceinline__ u64 mod_add(u64 a, u64 b, u64 n) {   // a, b < n
  const u64 s = a + b;
  return (s < a || s >= n) ? s - n : s;
}
__device__ __forceinline__ u64 mod_sub(u64 a, u64 b, u64 n) { return a >= b ? a - b : a - b + n; }

// Jacobi symbol (a/m), m odd, both below 2^32
__device__ __forceinline__ int jacobi32(u32 a, u32 m) {
  int t = 1;
  a %= m;
  while (a) {
    while (!(a & 1)) {
      a >>= 1;
      const u32 r = m & 7;
      if (r == 3 || r == 5) t = -t;
    }
    const u32 x = a;
    a = m;
    m = x;
    if ((a & 3) == 3 && (m & 3) == 3) t = -t;
    a %= m;
  }
  return m == 1 ? t : 0;
}

// (D/n) for odd |D| < 2^31 and odd n: reciprocity brings it to (n mod |D| / |D|)
__device__ __forceinline__ int jacobi_small(int D, u64 n) {
  const u32 a = (u32)(D < 0 ? -D : D);
  int t = (D < 0 && (n & 3) == 3) ? -1 : 1;           // (-1/n)
  if ((a & 3) == 3 && (n & 3) == 3) t = -t;            // (a/n) = (n/a) unless both are 3 mod 4
  const u32 hi = (u32)(n >> 32), lo = (u32)n;          // n mod a from 32-bit pieces
  const u32 r = (u32)((((u64)(hi % a) << 32) | lo) % a);
  return t * jacobi32(r, a);
}

__device__ __forceinline__ bool is_square64(u64 n) {
  u64 r = (u64)__builtin_sqrt((double)n);
  if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
  while (r * r > n) r--;
  while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) r++;
  return r * r == n;
}

// v * x mod n for a small v >= 0 (double and add), x < n
__device__ __forceinline__ u64 mul_small_mod(u32 v, u64 x, u64 n) {
  u64 r = 0;
  for (int b = 31 - __builtin_clz(v | 1); b >= 0; b--) {
    r = mod_add(r, r, n);
    if ((v >> b) & 1) r = mod_add(r, x, n);
  }
  return r;
}

// Montgomery form R mod n, R = 2^64
__device__ __forceinline__ u64 mont64_one(u64 n) {
  const int nb = 64 - __builtin_clzll(n);
  u64 one = 1ull << (nb - 1);
  for (int k = nb - 1; k < 64; k++) one = mont64_dbl(one, n);
  return one;
}

// true iff n is a strong Lucas probable prime with Selfridge's parameters (odd n >= 3)
__device__ __forceinline__ bool slprp_64(u64 n) {
  if (is_square64(n)) return false;                    // no D would have (D/n) = -1
  // For a non-square n the search ends within a few hundred steps; the bound only keeps
  // the loop finite by construction.
  int D = 5;
  for (;;) {
    const int j = jacobi_small(D, n);
    if (j < 0) break;
    if (j == 0 && (u64)(D < 0 ? -D : D) != n) return false;   // D shares a factor with n
    D = D > 0 ? -(D + 2) : 2 - D;
    if (D > 1000001 || D < -1000001) return false;
  }
  const int Q = (1 - D) / 4;                            // exact: D = 1 (mod 4)
  const u64 ninv = 0 - inv64(n);
  const u64 one = mont64_one(n);
  const u64 qa = mul_small_mod((u32)(Q < 0 ? -Q : Q), one, n);
  const u64 Qm = Q < 0 ? mod_sub(0, qa, n) : qa;        // Q in Montgomery form
  // n + 1 does not wrap: 2^64 - 1 is a multiple of 5 and was rejected by the search above
  const int s = __builtin_ctzll(n + 1);
  const u64 d = (n + 1) >> s;
  // k = 1: V_1 = P = 1, V_2 = P^2 - 2Q, Q^1
  u64 vk = one, vk1 = mod_sub(one, mod_add(Qm, Qm, n), n), qk = Qm;
  for (int i = 62 - __builtin_clzll(d); i >= 0; i--) {
    const bool bit = (d >> i) & 1;
    const u64 t = mod_sub(mont64_mul(vk, vk1, n, ninv), qk, n);     // V_(2k+1) = V_k V_(k+1) - P Q^k
    const u64 qq1 = mont64_mul(qk, Qm, n, ninv);
    const u64 x = bit ? vk1 : vk;
    const u64 qq = bit ? qq1 : qk;                                   // Q^(k + bit)
    const u64 sq = mod_sub(mont64_sqr(x, n, ninv), mod_add(qq, qq, n), n);       // V_(2(k + bit))
    qk = mont64_mul(qq, qk, n, ninv);                                // Q^(2k + bit)
    vk = bit ? t : sq;
    vk1 = bit ? sq : t;
  }
  // U_d = 0  <=>  2 V_(d+1) = V_d (mod n)
  if (mod_add(vk1, vk1, n) == vk || vk == 0) return true;
  for (int r = 1; r < s; r++) {
    vk = mod_sub(mont64_sqr(vk, n, ninv), mod_add(qk, qk, n), n);       // V_(2k) = V_k^2 - 2 Q^k
    if (vk == 0) return true;
    qk = mont64_sqr(qk, n, ninv);
  }
  return false;
}

// BPSW verdict bits: 1 = base-2 strong probable prime, 2 = strong Lucas probable prime,
// 4 = prime (both, for odd n >= 3; n = 2 is prime; other even n and n < 2 are not)
__device__ __forceinline__ u32 bpsw_flags(u64 n) {
  if (n < 3 || !(n & 1)) return n == 2 ? 4u : 0u;
  const u32 f = (sprp2_64(n) ? 1u : 0u) | (slprp_64(n) ? 2u : 0u);
  return f == 3u ? 7u : f;
}

// ---------------------------------------------------------------------------------
// order sieve for the SPRP-2 pass
// ---------------------------------------------------------------------------------

// x mod q for x < 2^64 by Barrett with minv = floor((2^64 - 1) / q)
__device__ __forceinline__ u32 mod_barrett(u64 x, u32 q, u64 minv) {
  const u64 t = __umul64hi(x, minv);
  u64 r = x - t * q;
  if (r >= q) r -= q;
  return (u32)r;
}

__host__ __device__ __forceinline__ u32 umin32(u32 x, u32 y) { return x < y ? x : y; }

// (-base) mod q for 512 <= q < 2^31 and a base below 2^53, from rq = 1/q rounded to nearest,
// with one-sided error.  base / q < 2^44, so base * rq is within 2^-9 of base / q (rq's
// relative error is at most 2^-53) and the FMA that subtracts 2^-7 rounds by at most 2^-9 more:
// its result lies in [base/q - 2^-7 - 2^-8, base/q - 2^-7 + 2^-8], below base / q and above
// base / q - 1.  So t = floor of it is floor(base / q) or one less, and r = base - t q (exact in
// the FMA, an integer in [0, 2q)) needs one correction; unsigned min(x, x - q) is x mod q for
// x < 2q.  Ten VALU instructions against twenty for the two-sided correction in doubles.
__host__ __device__ __forceinline__ u32 neg_mod_fp(double bd, u32 q, double rq) {
  const double t = __builtin_floor(__builtin_fma(bd, rq, -0.0078125));
  const u32 r2 = (u32)__builtin_fma(-t, (double)q, bd);
  const u32 r = umin32(r2, r2 - q);
  const u32 o = q - r;        // (0, q]
  return umin32(o, o - q);
}

// One prime of the order sieve.  A base-2 strong pseudoprime n (indeed any Fermat one) has
// 2^(n-1) = 1 (mod p) for each prime factor p, so ord_p(2) | n - 1.  With lcm(2, ord_p(2)) =
// s2, an odd multiple n = m p can be one only if m = u (mod s2), u = p^-1 mod s2 (p m = 1
// (mod s2)): every other odd multiple of p is ruled out.  mod = p s2 < 2^32 (p < 2^15) with
// its Barrett constant; rp = 1/p rounded to float.
struct PspOrder { u32 p, s2, u, mod; u64 minv; float rp; u32 pad; };

__host__ __device__ __forceinline__ u32 powmod32(u32 b, u32 x, u32 m) {
  u64 r = 1, y = b % m;
  for (; x; x >>= 1) {
    if (x & 1) r = r * y % m;
    y = y * y % m;
  }
  return (u32)r;
}

// The entry of an odd prime 3 < p < 2^15 (built on the host): ord_p(2) is p - 1 stripped of
// each prime factor f while 2^(L/f) = 1 (mod p).
__host__ __device__ __forceinline__ PspOrder psp_order_entry(u32 p) {
  u32 L = p - 1, x = p - 1;
  for (u32 f = 2; f <= x; f++) {
    if (x % f) continue;
    while (x % f == 0) x /= f;
    while (L % f == 0 && powmod32(2, L / f, p) == 1) L /= f;
  }
  PspOrder F;
  F.p = p;
  F.s2 = L % 2 ? 2 * L : L;
  F.u = 1;
  while ((u64)F.u * p % F.s2 != 1) F.u++;
  F.mod = p * F.s2;
  F.minv = 0xFFFFFFFFFFFFFFFFull / F.mod;
  F.rp = 1.0f / (float)p;
  F.pad = 0;
  return F;
}

// The odd numbers A, A + 2, ..., A + 2 (nodd - 1) (A odd, nodd <= 2^30) that the prime F
// rules out: ruled(i) for each, i its index.  The first odd multiple m p >= A comes from
// R = A mod (p s2): A - R is an even multiple of p, so m = ceil(R / p) made odd; then the
// walk keeps m mod s2 in a counter.
template <class G>
__device__ __forceinline__ void psp_order_walk(u64 A, const PspOrder& F, u32 nodd, G ruled) {
  const u32 R = mod_barrett(A, F.mod, F.minv);
  u32 q = (u32)((float)R * F.rp);              // floor(R / p) within a step or two
  while (q * F.p > R) q--;
  while ((q + 1) * F.p <= R) q++;
  u32 m = (q + (q * F.p != R)) | 1u;           // the first odd cofactor, <= s2 + 1
  u32 off = m * F.p - R;                       // A + off = (A - R) + m p, off < 2p
  u32 mm = m >= F.s2 ? m - F.s2 : m;           // m mod s2
  for (; off < 2u * nodd; off += 2u * F.p) {
    if (mm != F.u) ruled(off >> 1);
    mm += 2u;
    if (mm >= F.s2) mm -= F.s2;
  }
}

}  // namespace wf
