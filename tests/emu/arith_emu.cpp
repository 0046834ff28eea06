// Host build of mathpub_amd/csrc/wf_arith.h for CPU tests of the kernel arithmetic.
#define WF_HOST_EMU 1
#include "wf_arith.h"
#include <algorithm>

using namespace wf;

extern "C" {

// 2^(p-1) mod p^2 through the FP64 p-adic path (p < 2^50): r = hi:lo
void emu_fp64(uint64_t p, uint64_t* hi, uint64_t* lo) {
  const uint64_t e = p - 1;
  const int eb = 64 - __builtin_clzll(e);
  double a, b;
  fp64_pow2<5>(p, e, eb, a, b);
  const unsigned __int128 r = (unsigned __int128)(uint64_t)b * p + (uint64_t)a;
  *hi = (uint64_t)(r >> 64);
  *lo = (uint64_t)r;
}

// The kernel's finish of one ladder's raw digits: r (normalised, r[0] = high, r[1] = low)
// and f64_quotient's verdict: r[2] = 1 and r[3] = q when it takes the fast path, else 0, 0.
static void finish(uint64_t p, double a, double b, uint64_t* r) {
  uint64_t q = 0;
  const bool fast = f64_quotient(a, b, (double)p, q);
  F64Mod m;
  m.p = (double)p;
  m.ip = 1.0 / m.p;
  f64_normalise(a, b, m);
  const unsigned __int128 x = (unsigned __int128)(uint64_t)b * p + (uint64_t)a;
  r[0] = (uint64_t)(x >> 64);
  r[1] = (uint64_t)x;
  r[2] = fast;
  r[3] = fast ? q : 0;
}

// two ladders at once (the kernel's form), with a wavefront-wide exponent length ebits
// that may exceed either prime's own; r[0..3] for p0 and r[4..7] for p1 as finish() gives
void emu_fp64x2(uint64_t p0, uint64_t p1, int ebits, uint64_t* r, int big) {
  double a0, b0, a1, b1;
  if (big) fp64_pow2x2<5, true>((double)p0, p0 - 1, (double)p1, p1 - 1, ebits, a0, b0, a1, b1);
  else fp64_pow2x2<5>((double)p0, p0 - 1, (double)p1, p1 - 1, ebits, a0, b0, a1, b1);
  finish(p0, a0, b0, r);
  finish(p1, a1, b1, r + 4);
}

// the same through the L-limb Montgomery ladder; full = carry-keeping variant
void emu_mont(int L, int full, uint64_t p, int ebits, uint64_t* hi, uint64_t* lo) {
  const uint64_t e = p - 1;
  uint32_t x[4] = {0, 0, 0, 0};
  if (L == 2) { uint32_t y[2]; mont_pow2<2, false>(p, e, ebits, y); x[0] = y[0]; x[1] = y[1]; }
  else if (L == 3) { uint32_t y[3]; mont_pow2<3, false>(p, e, ebits, y); x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; }
  else if (!full) { mont_pow2<4, false>(p, e, ebits, x); }
  else { mont_pow2<4, true>(p, e, ebits, x); }
  *lo = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
  *hi = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
}

// any base b through the L-limb binary ladder (mont_powb)
void emu_montb(int L, int full, uint32_t b, uint64_t p, int ebits, uint64_t* hi, uint64_t* lo) {
  const uint64_t e = p - 1;
  uint32_t x[4] = {0, 0, 0, 0};
  if (L == 2) { uint32_t y[2]; mont_powb<2, false>(p, e, ebits, b, y); x[0] = y[0]; x[1] = y[1]; }
  else if (L == 3) { uint32_t y[3]; mont_powb<3, false>(p, e, ebits, b, y); x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; }
  else if (!full) { mont_powb<4, false>(p, e, ebits, b, x); }
  else { mont_powb<4, true>(p, e, ebits, b, x); }
  *lo = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
  *hi = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
}


// window and powers of a base, as the engine sets them up: the largest w <= 5 with
// b^(2^w - 1) < 2^32, pw[d] = b^d
static int base_table(uint32_t b, uint32_t* pw) {
  int w = 5;
  for (;;) {
    unsigned __int128 x = 1;
    bool ok = true;
    for (int k = 1; k < (1 << w); k++) { x *= b; if (x >> 32) { ok = false; break; } }
    if (ok || w == 1) break;
    w--;
  }
  uint64_t x = 1;
  for (int d = 0; d < (1 << w); d++) { pw[d] = (uint32_t)x; x *= b; }
  return w;
}

void emu_fp64c(uint32_t base, uint64_t p0, uint64_t p1, int ebits, uint64_t* r, int big) {
  uint32_t pw[32];
  const int w = base_table(base, pw);
  const double d0 = (double)p0, d1 = (double)p1;
  double a0, b0, a1, b1;
#define WF_POWC(B)                                                                              \
  switch (w) {                                                                                  \
    case 5: fp64_powc_x2<5, B>(d0, p0 - 1, d1, p1 - 1, ebits, pw, pw, a0, b0, a1, b1); break;        \
    case 4: fp64_powc_x2<4, B>(d0, p0 - 1, d1, p1 - 1, ebits, pw, pw, a0, b0, a1, b1); break;        \
    case 3: fp64_powc_x2<3, B>(d0, p0 - 1, d1, p1 - 1, ebits, pw, pw, a0, b0, a1, b1); break;        \
    case 2: fp64_powc_x2<2, B>(d0, p0 - 1, d1, p1 - 1, ebits, pw, pw, a0, b0, a1, b1); break;        \
    default: fp64_powc_x2<1, B>(d0, p0 - 1, d1, p1 - 1, ebits, pw, pw, a0, b0, a1, b1); break;       \
  }
  if (big) {
    WF_POWC(true)
  } else {
    WF_POWC(false)
  }
#undef WF_POWC
  finish(p0, a0, b0, r);
  finish(p1, a1, b1, r + 4);
}

void emu_m28c(uint32_t base, uint64_t p, int ebits, uint64_t* hi, uint64_t* lo) {
  uint32_t pw[32], x[4];
  const int w = base_table(base, pw);
  switch (w) {
    case 5: m28_powc<5>(p, p - 1, ebits, pw, x); break;
    case 4: m28_powc<4>(p, p - 1, ebits, pw, x); break;
    case 3: m28_powc<3>(p, p - 1, ebits, pw, x); break;
    case 2: m28_powc<2>(p, p - 1, ebits, pw, x); break;
    default: m28_powc<1>(p, p - 1, ebits, pw, x); break;
  }
  *lo = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
  *hi = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
}

int emu_sprp2(uint64_t n) { return sprp2_64(n) ? 1 : 0; }

int emu_psp_possible(uint64_t n) { return psp_possible(n) ? 1 : 0; }

// psp_possible from a tile's residues advanced by offsets, as k_psp forms them (n = A + o,
// o even and below 2^12): the number of n in A, A + 2, ..., A + 2 (count - 1) where the two
// forms disagree
uint32_t emu_psp_tile_mismatch(uint64_t A, uint32_t count) {
  const uint32_t r1 = (uint32_t)(A % PSP_M1), r2 = (uint32_t)(A % PSP_M2), r3 = (uint32_t)(A % PSP_M3),
                 r4 = (uint32_t)(A % PSP_M4), r5 = (uint32_t)(A % 97u);
  uint32_t bad = 0;
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t o = 2u * i;
    const uint64_t n = A + o;
    const uint32_t g1 = r1 + o, g2 = r2 + o, g3 = r3 + o, g4 = r4 + o;
    const bool t = psp_possible_res((uint32_t)n & 15u, g1 >= PSP_M1 ? g1 - PSP_M1 : g1, g2 >= PSP_M2 ? g2 - PSP_M2 : g2,
                                    g3 >= PSP_M3 ? g3 - PSP_M3 : g3, g4 >= PSP_M4 ? g4 - PSP_M4 : g4, (r5 + o) % 97u);
    bad += t != psp_possible(n);
  }
  return bad;
}

int emu_slprp(uint64_t n) { return slprp_64(n) ? 1 : 0; }


uint32_t emu_bpsw(uint64_t n) { return bpsw_flags(n); }

int emu_jacobi(int D, uint64_t n) { return jacobi_small(D, n); }

void emu_m28(uint64_t p, int ebits, uint64_t* hi, uint64_t* lo) {
  uint32_t x[4];
  m28_pow2(p, p - 1, ebits, x);
  *lo = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
  *hi = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
}

// One lazy squaring x^2 / R mod N on limbs as the ladder holds them (x[i] may exceed 28 bits)
void emu_m28_sqr(uint64_t p, const uint32_t* xin, uint32_t* xout) {
  M28Mod m;
  m28_setup(p, m);
  uint32_t x[5];
  for (int i = 0; i < 5; i++) x[i] = xin[i];
  m28_sqr(x, m.n, m.ninv);
  for (int i = 0; i < 5; i++) xout[i] = x[i];
}

// The order sieve of the SPRP-2 pass over the odd numbers A, A + 2, ..., A + 2 (nodd - 1):
// out[i] = 1 where some prime 101 <= p < pmax rules A + 2 i out (psp_order_entry and
// psp_order_walk, as k_psp runs them).
void emu_psp_order(uint64_t A, uint32_t pmax, uint32_t nodd, uint8_t* out) {
  for (uint32_t i = 0; i < nodd; i++) out[i] = 0;
  for (uint32_t p = 101; p < pmax; p += 2) {
    bool prime = true;
    for (uint32_t d = 3; d * d <= p; d += 2)
      if (p % d == 0) { prime = false; break; }
    if (!prime) continue;
    const PspOrder F = psp_order_entry(p);
    psp_order_walk(A, F, nodd, [&](uint32_t i) { out[i] = 1; });
  }
}

// Two bases on one prime (the composite-base mode's ladder pairs): window W = the smaller of
// the two bases' windows, r[0..3] for base A and r[4..7] for base B as finish() gives.
void emu_fp64c2(uint32_t baseA, uint32_t baseB, uint64_t p, int ebits, uint64_t* r, int big) {
  uint32_t pa[32], pb[32];
  const int w = std::min(base_table(baseA, pa), base_table(baseB, pb));
  const double d = (double)p;
  double a0, b0, a1, b1;
#define WF_POWC2(B)                                                                              \
  switch (w) {                                                                                   \
    case 5: fp64_powc_x2<5, B>(d, p - 1, d, p - 1, ebits, pa, pb, a0, b0, a1, b1); break;          \
    case 4: fp64_powc_x2<4, B>(d, p - 1, d, p - 1, ebits, pa, pb, a0, b0, a1, b1); break;          \
    case 3: fp64_powc_x2<3, B>(d, p - 1, d, p - 1, ebits, pa, pb, a0, b0, a1, b1); break;          \
    case 2: fp64_powc_x2<2, B>(d, p - 1, d, p - 1, ebits, pa, pb, a0, b0, a1, b1); break;          \
    default: fp64_powc_x2<1, B>(d, p - 1, d, p - 1, ebits, pa, pb, a0, b0, a1, b1); break;         \
  }
  if (big) {
    WF_POWC2(true)
  } else {
    WF_POWC2(false)
  }
#undef WF_POWC2
  finish(p, a0, b0, r);
  finish(p, a1, b1, r + 4);
}

// (-base) mod q by the sieve's one-sided FP64 placement (neg_mod_fp, rq = 1/q rounded to nearest)
uint32_t emu_neg_mod_fp(uint64_t base, uint32_t q) { return neg_mod_fp((double)base, q, 1.0 / (double)q); }

// n random (base < 2^53, 512 <= q < 2^31) pairs from a splitmix64 stream, half of them with
// base / q within 2^-6 of an integer on either side; returns how many differ from the exact
// (-base) mod q
uint64_t emu_neg_mod_fp_random(uint64_t n, uint64_t seed) {
  uint64_t bad = 0, z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  for (uint64_t k = 0; k < n; k++) {
    const int qb = 10 + (int)(next() % 22);                    // q of 10 to 31 bits
    uint32_t q = (uint32_t)(next() >> (64 - qb)) | (1u << (qb - 1));
    if (q < 512) q = 512;
    if (q >= 0x80000000u) q = 0x7FFFFFFFu;
    uint64_t base = next() >> 11;                               // below 2^53
    if (k & 1) {                                                // base / q near an integer
      const uint64_t m = base / q;
      const uint64_t d = next() % (q / 64 + 1);
      base = (next() & 2) ? m * q + d : (m ? m * q - 1 - d : d);
      if (base >= (1ull << 53)) base = (1ull << 53) - 1 - d;
    }
    const uint32_t want = (uint32_t)((q - base % q) % q);
    if (emu_neg_mod_fp(base, q) != want) bad++;
  }
  return bad;
}

// 1/p by f64_recip_near from the base's reciprocal (p = base + f)
double emu_recip_near(uint64_t base, uint32_t f) {
  const double pd = (double)(base + f);
  return f64_recip_near(1.0 / (double)base, (double)f, pd);
}

// the kernel's two ladders above 2^40: p_k = base + f_k, 1/p_k from f64_recip_near
void emu_fp64x2_near(uint64_t base, uint32_t f0, uint32_t f1, int ebits, uint64_t* r) {
  const double ib = 1.0 / (double)base;
  const uint64_t p0 = base + f0, p1 = base + f1;
  const double pd0 = (double)p0, pd1 = (double)p1;
  double a0, b0, a1, b1;
  fp64_pow2x2_r<5, true>(pd0, f64_recip_near(ib, (double)f0, pd0), p0 - 1, pd1, f64_recip_near(ib, (double)f1, pd1),
                         p1 - 1, ebits, a0, b0, a1, b1);
  finish(p0, a0, b0, r);
  finish(p1, a1, b1, r + 4);
}

}  // extern "C"
