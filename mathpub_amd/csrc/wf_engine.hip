// wf_engine.hip -- the MI355X (gfx950) Wieferich search engine behind include/wf.h.
//
// Device pipeline of one pass over [lo, hi) (the work of scan_chunk, A001220/wieferich.c:
// 1129-1177, with the prime source of :1138-1141 moved onto the GPU):
//
//   k_sieve   one workgroup per wheel-30 segment held in LDS (bit c of byte k stands for
//             base + 30 k + R30[c]).  Small sieving primes are striped over the lanes of a
//             wavefront, the others take one lane each; marks are LDS atomic ORs.  The
//             surviving bits are extracted with popcount + a wavefront prefix scan and
//             appended to an HBM prime list (one atomic per wavefront).
//   k_test_*  one prime per lane: r = 2^(p-1) mod p^2 (wf_arith.h), q, the FLT check,
//             mix64 checksum, sum of q, FOUND / NEAR / FLT events; wavefront reductions
//             and one 64-bit atomic per wavefront and counter.
//
// The host side keeps the reference's conventions (wf.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "wf.h"
#include "wf_arith.h"

using namespace wf;
typedef int64_t i64;

#define GOLD 0x9E3779B97F4A7C15ULL   // wieferich.c:240

// ---------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------

// splitmix64 finalizer (wieferich.c:452-458)
__device__ __forceinline__ u64 d_mix64(u64 z) {
  z += GOLD;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ u64 wave_sum(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ u32 wave_incl_scan(u32 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u32 t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// x mod q for x < 2^64 by Barrett with minv = floor((2^64 - 1) / q)
__device__ __forceinline__ u32 mod_barrett(u64 x, u32 q, u64 minv) {
  const u64 t = __umul64hi(x, minv);
  u64 r = x - t * q;
  if (r >= q) r -= q;
  return (u32)r;
}

__constant__ u32 c_R30[8] = {1, 7, 11, 13, 17, 19, 23, 29};
// c_M30[qi][o mod 30][c]: the m in [0, 30) with o + m q = R30[c] (mod 30), qi = the class
// of q mod 30 (an index into R30).
__constant__ unsigned char c_M30[8 * 30 * 8];
// c_QI[q mod 30] = index of q mod 30 in R30 (q coprime to 30)
__constant__ unsigned char c_QI[30];

// Byte index j (relative to the segment base) of the first mark of class c for a prime q
// whose first multiple at or after max(base, q^2) lies at offset o (numbers) from base.
__device__ __forceinline__ u64 class_start(u64 o, u32 q, int qi, int c) {
  const u32 om = (u32)(o % 30);
  const u32 m = c_M30[(qi * 30 + om) * 8 + c];
  return (o + (u64)m * q - c_R30[c]) / 30;
}

// ---------------------------------------------------------------------------------
// sieve
// ---------------------------------------------------------------------------------

struct SegDesc {
  u64 base;       // multiple of 30
  u32 nbytes;     // wheel bytes (a multiple of 4, <= the LDS segment)
  u32 rid;        // range index
  u32 lo_off;     // valid numbers are base + [lo_off, hi_off)
  u32 hi_off;
  u32 nq;         // sieving primes with q^2 <= the segment's last number
  u32 pad;
};

struct DevAcc {     // per range; every field is a 64-bit sum
  unsigned long long primes, tests, near, flt, cks, sumq, nev, npsp;
};

struct DevEvent { u32 type, rid; u64 p, q, r_hi, r_lo; };

// first multiple offset of q at or after max(base, q^2), relative to base; returns false
// if it lies at or beyond 'span' numbers
__device__ __forceinline__ bool first_offset(u64 base, u32 q, u64 minv, u64 span, u64& o) {
  const u64 qq = (u64)q * q;
  if (qq >= base) {
    o = qq - base;
  } else {
    const u32 r = mod_barrett(base, q, minv);
    o = r ? q - r : 0;
  }
  return o < span;
}

__global__ void __launch_bounds__(1024) k_sieve(const SegDesc* __restrict__ segs, u32 nseg,
                                                const u32* __restrict__ sp, const u64* __restrict__ spinv,
                                                u32 tierA,
                                                u64* __restrict__ out_p, u32* __restrict__ out_rid,
                                                unsigned long long* __restrict__ ctr, u64 cap,
                                                DevAcc* __restrict__ acc) {
  extern __shared__ u32 seg[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  for (u32 s = blockIdx.x; s < nseg; s += gridDim.x) {
    const SegDesc d = segs[s];
    const u32 nwords = d.nbytes >> 2;
    const u64 span = 30ull * d.nbytes;
    for (u32 w = threadIdx.x; w < nwords; w += blockDim.x) seg[w] = 0;
    __syncthreads();
    // tier A: one wavefront per prime, lanes stride over the marks of each class
    for (u32 i = wave; i < min(tierA, d.nq); i += nwaves) {
      const u32 q = sp[i];
      u64 o;
      if (!first_offset(d.base, q, spinv[i], span, o)) continue;
      const int qi = c_QI[q % 30];
      for (int c = 0; c < 8; c++) {
        const u64 j0 = class_start(o, q, qi, c);
        for (u64 j = j0 + (u64)lane * q; j < d.nbytes; j += 64ull * q)
          atomicOr(&seg[j >> 2], 1u << ((((u32)j & 3) << 3) | c));
      }
    }
    // tier B: one lane per prime
    for (u32 i = tierA + threadIdx.x; i < d.nq; i += blockDim.x) {
      const u32 q = sp[i];
      u64 o;
      if (!first_offset(d.base, q, spinv[i], span, o)) continue;
      const int qi = c_QI[q % 30];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        for (u64 j = class_start(o, q, qi, c); j < d.nbytes; j += q)
          atomicOr(&seg[j >> 2], 1u << ((((u32)j & 3) << 3) | c));
      }
    }
    __syncthreads();
    // extraction: survivors of each word within [lo_off, hi_off)
    u32 cnt = 0;
    for (u32 w = threadIdx.x; w < nwords; w += blockDim.x) {
      u32 v = ~seg[w];
      const u32 w0 = 120u * w;  // offset of the word's first byte
      if (w0 < d.lo_off || w0 + 120u > d.hi_off) {
        u32 m = 0;
        for (int b = 0; b < 32; b++) {
          const u32 off = w0 + 30u * (b >> 3) + c_R30[b & 7];
          if (off >= d.lo_off && off < d.hi_off) m |= 1u << b;
        }
        v &= m;
      }
      cnt += __popc(v);
    }
    // the primes 3 and 5 are not on the wheel
    u32 extra = 0;
    if (d.base == 0 && threadIdx.x == 0) {
      if (3 >= d.lo_off && 3 < d.hi_off) extra |= 1;
      if (5 >= d.lo_off && 5 < d.hi_off) extra |= 2;
      cnt += __popc(extra);
    }
    const u32 incl = wave_incl_scan(cnt);
    const u32 tot = __shfl(incl, 63, 64);
    unsigned long long wbase = 0;
    if (lane == 63 && tot) {
      wbase = atomicAdd(&ctr[0], (unsigned long long)tot);
      atomicAdd(&acc[d.rid].primes, (unsigned long long)tot);
    }
    wbase = __shfl(wbase, 63, 64);
    u64 pos = wbase + incl - cnt;
    if (extra & 1) { if (pos < cap) { out_p[pos] = 3; out_rid[pos] = d.rid; } pos++; }
    if (extra & 2) { if (pos < cap) { out_p[pos] = 5; out_rid[pos] = d.rid; } pos++; }
    for (u32 w = threadIdx.x; w < nwords; w += blockDim.x) {
      u32 v = ~seg[w];
      const u32 w0 = 120u * w;
      if (w0 < d.lo_off || w0 + 120u > d.hi_off) {
        u32 m = 0;
        for (int b = 0; b < 32; b++) {
          const u32 off = w0 + 30u * (b >> 3) + c_R30[b & 7];
          if (off >= d.lo_off && off < d.hi_off) m |= 1u << b;
        }
        v &= m;
      }
      while (v) {
        const int b = __ffs(v) - 1;
        v &= v - 1;
        if (pos < cap) {
          out_p[pos] = d.base + w0 + 30u * (b >> 3) + c_R30[b & 7];
          out_rid[pos] = d.rid;
        }
        pos++;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------
// tests
// ---------------------------------------------------------------------------------

struct TestOut {
  DevAcc* acc;
  DevEvent* ev;
  u64 ev_cap;
  unsigned long long* nev;   // total events (device counter)
  i64 near_T;
};

__device__ __forceinline__ void push_event(const TestOut& o, u32 type, u32 rid, u64 p, u64 q, u64 rhi, u64 rlo) {
  const unsigned long long k = atomicAdd(o.nev, 1ull);
  atomicAdd(&o.acc[rid].nev, 1ull);
  if (k < o.ev_cap) {
    DevEvent e;
    e.type = type; e.rid = rid; e.p = p; e.q = q; e.r_hi = rhi; e.r_lo = rlo;
    o.ev[k] = e;
  }
}

// Account one tested prime exactly as finish_lanes (wieferich.c:1051-1075) and add the
// lane's terms to the wavefront sums.  r = r_hi:r_lo, q = (r - 1) p^-1 mod 2^64.
struct LaneSums { u64 tests, flt, near, cks, sumq; };

__device__ __forceinline__ void account(const TestOut& o, u32 rid, u64 p, u64 rhi, u64 rlo, LaneSums& s) {
  const u64 pinv = inv64(p);
  const u64 q = (rlo - 1) * pinv;
  s.tests++;
  if (q >= p) {
    s.flt++;
    push_event(o, WF_EV_FLT, rid, p, q, rhi, rlo);
    return;
  }
  s.cks += d_mix64(q ^ (p * GOLD));
  s.sumq += q;
  if (q == 0) {
    push_event(o, WF_EV_FOUND, rid, p, q, rhi, rlo);
  } else if (o.near_T > 0) {
    const i64 A = q <= p / 2 ? (i64)q : -(i64)(p - q);
    if ((A < 0 ? -A : A) <= o.near_T) {
      s.near++;
      push_event(o, WF_EV_NEAR, rid, p, q, rhi, rlo);
    }
  }
}

// Wavefront reduction of the lane sums into the range's accumulator; lanes of one
// wavefront usually share a range, otherwise each lane adds its own terms.
__device__ __forceinline__ void flush_sums(const TestOut& o, u32 rid, bool active, const LaneSums& s) {
  const int lane = threadIdx.x & 63;
  const u32 r0 = __shfl(rid, 0, 64);
  const bool uniform = __all(!active || rid == r0);
  if (uniform) {
    const u64 t = wave_sum(s.tests), f = wave_sum(s.flt), n = wave_sum(s.near);
    const u64 c = wave_sum(s.cks), q = wave_sum(s.sumq);
    if (lane == 0) {
      DevAcc* a = &o.acc[r0];
      if (t) atomicAdd(&a->tests, t);
      if (f) atomicAdd(&a->flt, f);
      if (n) atomicAdd(&a->near, n);
      atomicAdd(&a->cks, c);
      atomicAdd(&a->sumq, q);
    }
  } else if (active) {
    DevAcc* a = &o.acc[rid];
    atomicAdd(&a->tests, s.tests);
    if (s.flt) atomicAdd(&a->flt, s.flt);
    if (s.near) atomicAdd(&a->near, s.near);
    atomicAdd(&a->cks, s.cks);
    atomicAdd(&a->sumq, s.sumq);
  }
}

// 128-bit r = a + b p from the normalised p-adic digits
__device__ __forceinline__ void padic_to_r(u64 p, double a, double b, u64& rhi, u64& rlo) {
  const u64 ai = (u64)a, bi = (u64)b;
  const u64 lo = bi * p;
  const u64 hi = __umul64hi(bi, p);
  rlo = lo + ai;
  rhi = hi + (rlo < lo);
}

template <int W>
__global__ void __launch_bounds__(256) k_test_f64(const u64* __restrict__ primes, const u32* __restrict__ rid,
                                                  const unsigned long long* __restrict__ count, u64 cap,
                                                  TestOut o) {
  const u64 n = min((u64)*count, cap);   // the sieve never writes past cap
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 base = (u64)blockIdx.x * blockDim.x; base < n; base += stride) {
    const u64 i = base + threadIdx.x;
    const bool active = i < n;
    const u64 p = active ? primes[i] : 3;
    const u32 r = active ? rid[i] : 0;
    const u64 e = p - 1;
    const int eb = wave_max(64 - __builtin_clzll(e));
    double a, b;
    fp64_pow2<W>(p, e, eb, a, b);
    LaneSums s = {0, 0, 0, 0, 0};
    if (active) {
      u64 rhi, rlo;
      padic_to_r(p, a, b, rhi, rlo);
      account(o, r, p, rhi, rlo, s);
    }
    flush_sums(o, r, active, s);
  }
}

template <int L, bool FULL>
__global__ void __launch_bounds__(256) k_test_mont(const u64* __restrict__ primes, const u32* __restrict__ rid,
                                                   const unsigned long long* __restrict__ count, u64 cap,
                                                   TestOut o) {
  const u64 n = min((u64)*count, cap);
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 base = (u64)blockIdx.x * blockDim.x; base < n; base += stride) {
    const u64 i = base + threadIdx.x;
    const bool active = i < n;
    const u64 p = active ? primes[i] : (FULL ? 0xFFFFFFFFFFFFFFC5ull : (L == 2 ? 3ull : (L == 3 ? 2147483659ull : 140737488355333ull)));
    const u32 r = active ? rid[i] : 0;
    const u64 e = p - 1;
    const int eb = wave_max(64 - __builtin_clzll(e));
    u32 x[L];
    mont_pow2<L, FULL>(p, e, eb, x);
    LaneSums s = {0, 0, 0, 0, 0};
    if (active) {
      const u64 rlo = (u64)x[0] | ((u64)x[1] << 32);
      const u64 rhi = (L > 2 ? (u64)x[2] : 0) | (L > 3 ? ((u64)x[3] << 32) : 0);
      account(o, r, p, rhi, rlo, s);
    }
    flush_sums(o, r, active, s);
  }
}

// Quotients of arbitrary odd p (check / selftest): per-lane choice of the path.
__global__ void k_quot(const u64* __restrict__ p_in, u64* __restrict__ q_out, u64* __restrict__ rhi_out,
                       u64* __restrict__ rlo_out, u64 n) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 p = p_in[i];
  const u64 e = p - 1;
  const int eb = 64 - __builtin_clzll(e);
  u64 rhi, rlo;
  if (p < (1ull << 50)) {
    double a, b;
    fp64_pow2<5>(p, e, eb, a, b);
    padic_to_r(p, a, b, rhi, rlo);
  } else if (p < (1ull << 63)) {
    u32 x[4];
    mont_pow2<4, false>(p, e, eb, x);
    rlo = (u64)x[0] | ((u64)x[1] << 32);
    rhi = (u64)x[2] | ((u64)x[3] << 32);
  } else {
    u32 x[4];
    mont_pow2<4, true>(p, e, eb, x);
    rlo = (u64)x[0] | ((u64)x[1] << 32);
    rhi = (u64)x[2] | ((u64)x[3] << 32);
  }
  q_out[i] = (rlo - 1) * inv64(p);
  rhi_out[i] = rhi;
  rlo_out[i] = rlo;
}

// ---------------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------------

static const u32 R30H[8] = {1, 7, 11, 13, 17, 19, 23, 29};

struct wf_engine {
  int dev = 0;
  hipStream_t st = nullptr;
  hipDeviceProp_t prop;
  wf_config cfg;
  int nb = 1;
  u32 bases[WF_MAXB] = {2};
  // sieving primes (odd primes >= 7, ascending)
  std::vector<u32> h_sp;
  u64 sp_lim = 0;
  u32* d_sp = nullptr;
  u64* d_spinv = nullptr;
  u32 tierA = 0;
  // work buffers
  SegDesc* d_segs = nullptr;
  u64 seg_cap = 0;
  u64* d_primes = nullptr;
  u32* d_rid = nullptr;
  u64 prime_cap = 0, rid_cap = 0;
  DevAcc* d_acc = nullptr;
  u64 acc_cap = 0;
  DevEvent* d_ev = nullptr;
  u64 dev_ev_cap = 0;
  unsigned long long* d_ctr = nullptr;   // [0] primes appended, [1] events
  hipEvent_t ev[4];
  wf_timing last;
};

#define HIPCHK(x)                                  \
  do {                                             \
    if ((x) != hipSuccess) return WF_E_DEVICE;     \
  } while (0)

extern "C" const char* wf_strerror(int rc) {
  switch (rc) {
    case WF_OK: return "ok";
    case WF_E_ARG: return "bad argument";
    case WF_E_DEVICE: return "GPU not available or HIP runtime error";
    case WF_E_NOMEM: return "out of memory";
    case WF_E_UNSUPPORTED: return "not supported by the GPU engine (base 2 only)";
    case WF_E_OVERFLOW: return "internal buffer bound exceeded";
    case WF_E_BUSY: return "serving: no free slot";
    default: return "unknown error";
  }
}

static u64 isqrt_u64(u64 n) {
  u64 r = (u64)std::sqrt((double)n);
  while (r > 0 && (unsigned __int128)r * r > n) r--;
  while ((unsigned __int128)(r + 1) * (r + 1) <= n) r++;
  return r;
}

// Odd primes 7 <= q <= lim on the host (segmented odd-only sieve), lim <= 2^32.
static void host_primes(u64 lim, std::vector<u32>& out) {
  out.clear();
  const u64 root = isqrt_u64(lim);
  std::vector<uint8_t> small(root + 1, 1);
  std::vector<u32> base;
  for (u64 i = 2; i <= root; i++) {
    if (!small[i]) continue;
    if (i > 2) base.push_back((u32)i);
    for (u64 k = i * i; k <= root; k += i) small[k] = 0;
  }
  const u64 SEGN = 1u << 22;  // odd numbers per segment
  std::vector<uint8_t> comp(SEGN);
  for (u64 lo = 1; lo <= lim; lo += 2 * SEGN) {  // odd numbers lo, lo + 2, ...
    const u64 hi = std::min(lim + 1, lo + 2 * SEGN);
    const u64 cnt = (hi - lo + 1) / 2;
    std::fill(comp.begin(), comp.begin() + cnt, 0);
    for (u32 q : base) {
      const u64 qq = (u64)q * q;
      if (qq >= hi) break;
      u64 f = qq >= lo ? qq : lo + (q - lo % q) % q;
      if (!(f & 1)) f += q;
      for (u64 k = (f - lo) / 2; k < cnt; k += q) comp[k] = 1;
    }
    for (u64 k = 0; k < cnt; k++) {
      const u64 n = lo + 2 * k;
      if (!comp[k] && n >= 7 && n <= lim) out.push_back((u32)n);
    }
  }
}

static int ensure_sieving_primes(wf_engine* e, u64 lim) {
  if (lim <= e->sp_lim) return WF_OK;
  u64 want = std::max<u64>(lim, std::min<u64>(std::max<u64>(2 * e->sp_lim, 1u << 20), 0xFFFFFFFFull));
  if (want > 0xFFFFFFFFull) want = 0xFFFFFFFFull;
  host_primes(want, e->h_sp);
  std::vector<u64> inv(e->h_sp.size());
  for (size_t i = 0; i < e->h_sp.size(); i++) inv[i] = 0xFFFFFFFFFFFFFFFFull / e->h_sp[i];
  if (e->d_sp) hipFree(e->d_sp);
  if (e->d_spinv) hipFree(e->d_spinv);
  e->d_sp = nullptr;
  e->d_spinv = nullptr;
  const size_t n = std::max<size_t>(e->h_sp.size(), 1);
  if (hipMalloc(&e->d_sp, n * sizeof(u32)) != hipSuccess) return WF_E_NOMEM;
  if (hipMalloc(&e->d_spinv, n * sizeof(u64)) != hipSuccess) return WF_E_NOMEM;
  HIPCHK(hipMemcpy(e->d_sp, e->h_sp.data(), e->h_sp.size() * sizeof(u32), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->d_spinv, inv.data(), inv.size() * sizeof(u64), hipMemcpyHostToDevice));
  e->sp_lim = want;
  // tier A: primes below 64 * 8 = 512 are striped over a wavefront
  e->tierA = (u32)(std::lower_bound(e->h_sp.begin(), e->h_sp.end(), 512u) - e->h_sp.begin());
  return WF_OK;
}

template <typename T>
static int grow(T*& ptr, u64& cap, u64 need) {
  if (need <= cap) return WF_OK;
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  u64 n = std::max<u64>(need, cap + cap / 2);
  if (hipMalloc(&ptr, n * sizeof(T)) != hipSuccess) {
    cap = 0;
    return WF_E_NOMEM;
  }
  cap = n;
  return WF_OK;
}

extern "C" int wf_open(int device, const wf_config* cfg, wf_engine** out) {
  if (!out) return WF_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) return WF_E_DEVICE;
  HIPCHK(hipSetDevice(device));
  wf_engine* e = new wf_engine();
  e->dev = device;
  if (hipGetDeviceProperties(&e->prop, device) != hipSuccess) { delete e; return WF_E_DEVICE; }
  memset(&e->cfg, 0, sizeof e->cfg);
  if (cfg) e->cfg = *cfg;
  if (!e->cfg.seg_bytes) e->cfg.seg_bytes = 32768;
  if (!e->cfg.sieve_block) e->cfg.sieve_block = 512;
  if (!e->cfg.test_block) e->cfg.test_block = 256;
  if (!e->cfg.chunk) e->cfg.chunk = 10000000000ull;
  if (e->cfg.seg_bytes < 4096 || (e->cfg.seg_bytes & (e->cfg.seg_bytes - 1)) ||
      e->cfg.seg_bytes > e->prop.sharedMemPerBlock || e->cfg.sieve_block < 64 || e->cfg.sieve_block > 1024 ||
      (e->cfg.sieve_block & 63) || e->cfg.test_block != 256) {
    delete e;
    return WF_E_ARG;
  }
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) { delete e; return WF_E_DEVICE; }
  for (int i = 0; i < 4; i++) hipEventCreate(&e->ev[i]);
  // wheel tables
  unsigned char m30[8 * 30 * 8], qi[30];
  memset(qi, 0, sizeof qi);
  for (int k = 0; k < 8; k++) qi[R30H[k]] = (unsigned char)k;
  for (int k = 0; k < 8; k++) {
    const u32 q = R30H[k];
    for (u32 o = 0; o < 30; o++)
      for (int c = 0; c < 8; c++) {
        u32 m = 0;
        while ((o + m * q) % 30 != R30H[c]) m++;
        m30[(k * 30 + o) * 8 + c] = (unsigned char)m;
      }
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_M30), m30, sizeof m30) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_QI), qi, sizeof qi) != hipSuccess) {
    wf_close(e);
    return WF_E_DEVICE;
  }
  if (hipMalloc(&e->d_ctr, 8 * sizeof(unsigned long long)) != hipSuccess) { wf_close(e); return WF_E_NOMEM; }
  memset(&e->last, 0, sizeof e->last);
  *out = e;
  return WF_OK;
}

extern "C" void wf_close(wf_engine* e) {
  if (!e) return;
  hipSetDevice(e->dev);
  if (e->st) hipStreamSynchronize(e->st);
  hipFree(e->d_sp);
  hipFree(e->d_spinv);
  hipFree(e->d_segs);
  hipFree(e->d_primes);
  hipFree(e->d_rid);
  hipFree(e->d_acc);
  hipFree(e->d_ev);
  hipFree(e->d_ctr);
  for (int i = 0; i < 4; i++) hipEventDestroy(e->ev[i]);
  if (e->st) hipStreamDestroy(e->st);
  delete e;
}

extern "C" int wf_info(wf_engine* e, char* buf, int len) {
  if (!e || !buf || len <= 0) return WF_E_ARG;
  snprintf(buf, (size_t)len, "%s (%s), %d CUs, LDS segment %u B, %s", e->prop.name, e->prop.gcnArchName,
           e->prop.multiProcessorCount, e->cfg.seg_bytes,
           e->cfg.arith == 1 ? "32-bit Montgomery" : "FP64 p-adic below 2^50, 32-bit Montgomery above");
  return WF_OK;
}

extern "C" int wf_set_bases(wf_engine* e, const u32* bases, int nb) {
  if (!e || !bases || nb < 1 || nb > WF_MAXB) return WF_E_ARG;
  for (int i = 0; i < nb; i++)
    if (bases[i] < 2 || (i && bases[i] <= bases[i - 1])) return WF_E_ARG;
  if (nb != 1 || bases[0] != 2) return WF_E_UNSUPPORTED;
  e->nb = nb;
  memcpy(e->bases, bases, (size_t)nb * sizeof(u32));
  return WF_OK;
}

// A sub-range of one arithmetic class: [lo, hi) (hi == 0 means 2^64) with its kernel.
enum Arith { A_F64 = 0, A_M2, A_M3, A_M4, A_FULL };
struct Piece { u64 lo, hi; u32 rid; int arith; };

// Split [lo, hi) at the class boundaries of the arithmetic paths.
static void split_classes(const wf_engine* e, u64 lo, u64 hi, u32 rid, std::vector<Piece>& out) {
  std::vector<std::pair<u64, int>> cuts;   // (upper bound of the class, class)
  if (e->cfg.arith == 1) {
    cuts = {{1ull << 31, A_M2}, {1ull << 47, A_M3}, {1ull << 63, A_M4}, {0, A_FULL}};
  } else {
    cuts = {{1ull << 50, A_F64}, {1ull << 63, A_M4}, {0, A_FULL}};
  }
  u64 cur = lo;
  for (auto& c : cuts) {
    const u64 top = c.first;  // 0 = 2^64
    if (top != 0 && cur >= top) continue;
    const u64 end = (hi == 0) ? top : (top == 0 ? hi : std::min(hi, top));
    if (end != 0 && end <= cur) break;
    out.push_back({cur, end, rid, c.second});
    if (end == hi) break;
    cur = end;
  }
}

// Numbers in [lo, hi) (hi == 0: 2^64) as a double, for sizing.
static double span_d(u64 lo, u64 hi) { return hi ? (double)(hi - lo) : 18446744073709551616.0 - (double)lo; }

// Upper bound on the primes in an interval of y numbers (Montgomery-Vaughan:
// pi(x + y) - pi(x) < 2 y / ln y for y > 1), plus slack for tiny y.
static u64 prime_bound(double y) { return y < 100 ? 64 : (u64)(2.0 * y / std::log(y)) + 64; }

// Build the segment descriptors of [lo, hi) into segs.
static void build_segments(const wf_engine* e, u64 lo, u64 hi, u32 rid, std::vector<SegDesc>& segs) {
  const u32 SB = e->cfg.seg_bytes;
  const u64 span = 30ull * SB;
  u64 base = lo - lo % 30;
  for (;;) {
    SegDesc d;
    d.base = base;
    d.rid = rid;
    // numbers of this segment: [base, base + span) capped at hi (or 2^64)
    const u64 room = 0 - base;  // 2^64 - base (0 if base == 0: unlimited)
    u64 len = span;
    if (base != 0 && room < len) len = room;
    if (hi != 0 && hi - base < len) len = hi - base;
    const u32 nbytes = (u32)((len + 29) / 30);
    d.nbytes = (nbytes + 3) & ~3u;
    d.lo_off = lo > base ? (u32)(lo - base) : 0;
    if (base == 0 && d.lo_off < 2) d.lo_off = 2;   // 1 is not prime
    d.hi_off = (u32)len;
    // sieving primes needed: q^2 <= last number of the segment
    const u64 last = base + (len - 1);
    const u64 r = isqrt_u64(last);
    d.nq = (u32)(std::upper_bound(e->h_sp.begin(), e->h_sp.end(), (u32)std::min<u64>(r, 0xFFFFFFFFull)) -
                 e->h_sp.begin());
    d.pad = 0;
    segs.push_back(d);
    if (len < span) break;
    base += span;
    if (hi != 0 && base >= hi) break;
    if (base == 0) break;
  }
}

// Sieve and test a list of pieces (all of one arithmetic class) whose total prime count
// fits the prime buffer; accumulators indexed by Piece::rid.
static int run_pieces(wf_engine* e, const std::vector<Piece>& pieces, int arith, TestOut& to, float& sieve_ms,
                      float& test_ms) {
  std::vector<SegDesc> segs;
  for (const Piece& p : pieces) build_segments(e, p.lo, p.hi, p.rid, segs);
  if (segs.empty()) return WF_OK;
  int rc;
  if ((rc = grow(e->d_segs, e->seg_cap, segs.size())) != WF_OK) return rc;
  u64 bound = 0;
  for (const Piece& p : pieces) bound += prime_bound(span_d(p.lo, p.hi));
  if ((rc = grow(e->d_primes, e->prime_cap, bound)) != WF_OK) return rc;
  if ((rc = grow(e->d_rid, e->rid_cap, e->prime_cap)) != WF_OK) return rc;
  HIPCHK(hipMemcpyAsync(e->d_segs, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemsetAsync(e->d_ctr, 0, sizeof(unsigned long long), e->st));
  const u32 SB = e->cfg.seg_bytes;
  const int cus = e->prop.multiProcessorCount;
  const u32 nseg = (u32)segs.size();
  const u32 grid = std::min<u32>(nseg, (u32)cus * 4);
  HIPCHK(hipEventRecord(e->ev[0], e->st));
  hipLaunchKernelGGL(k_sieve, dim3(grid), dim3(e->cfg.sieve_block), SB, e->st, e->d_segs, nseg, e->d_sp,
                     e->d_spinv, e->tierA, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to.acc);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e->ev[1], e->st));
  const u32 tgrid = (u32)std::min<u64>((bound + 255) / 256, (u64)cus * 64);
  switch (arith) {
    case A_F64: hipLaunchKernelGGL(k_test_f64<5>, dim3(tgrid), dim3(256), 0, e->st, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to); break;
    case A_M2: hipLaunchKernelGGL((k_test_mont<2, false>), dim3(tgrid), dim3(256), 0, e->st, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to); break;
    case A_M3: hipLaunchKernelGGL((k_test_mont<3, false>), dim3(tgrid), dim3(256), 0, e->st, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to); break;
    case A_M4: hipLaunchKernelGGL((k_test_mont<4, false>), dim3(tgrid), dim3(256), 0, e->st, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to); break;
    default: hipLaunchKernelGGL((k_test_mont<4, true>), dim3(tgrid), dim3(256), 0, e->st, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, to); break;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(e->ev[2], e->st));
  // check the prime-buffer bound (a proven bound; a violation is a bug)
  unsigned long long appended = 0;
  HIPCHK(hipMemcpyAsync(&appended, e->d_ctr, sizeof appended, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  if (appended > e->prime_cap) return WF_E_OVERFLOW;
  float a = 0, b = 0;
  hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
  hipEventElapsedTime(&b, e->ev[1], e->ev[2]);
  sieve_ms += a;
  test_ms += b;
  e->last.tested += appended;
  e->last.launches += 2;
  return WF_OK;
}

static bool cmp_dev_event(const DevEvent& a, const DevEvent& b) {
  if (a.p != b.p) return a.p < b.p;
  return a.type < b.type;
}

// Scan n ranges; per-range totals into res[0..n), events into ev.
static int scan_ranges(wf_engine* e, const wf_range* r, u32 n, i64 near_T, wf_result* res, wf_event* ev,
                       u64 ev_cap) {
  if (!e || !r || !res || n == 0) return WF_E_ARG;
  if (e->nb != 1 || e->bases[0] != 2) return WF_E_UNSUPPORTED;
  HIPCHK(hipSetDevice(e->dev));
  u64 maxtop = 0;
  for (u32 i = 0; i < n; i++) {
    if (r[i].hi != 0 && r[i].hi <= r[i].lo) return WF_E_ARG;
    const u64 top = r[i].hi ? r[i].hi - 1 : 0xFFFFFFFFFFFFFFFFull;
    maxtop = std::max(maxtop, top);
  }
  int rc;
  if ((rc = ensure_sieving_primes(e, std::max<u64>(isqrt_u64(maxtop) + 1, 64))) != WF_OK) return rc;
  if ((rc = grow(e->d_acc, e->acc_cap, n)) != WF_OK) return rc;
  const u64 want_ev = std::max<u64>(ev_cap, 1 << 16);
  if ((rc = grow(e->d_ev, e->dev_ev_cap, want_ev)) != WF_OK) return rc;
  HIPCHK(hipMemsetAsync(e->d_acc, 0, n * sizeof(DevAcc), e->st));
  HIPCHK(hipMemsetAsync(e->d_ctr + 1, 0, sizeof(unsigned long long), e->st));
  TestOut to;
  to.acc = e->d_acc;
  to.ev = e->d_ev;
  to.ev_cap = e->dev_ev_cap;
  to.nev = e->d_ctr + 1;
  to.near_T = near_T;
  memset(&e->last, 0, sizeof e->last);
  // pieces by arithmetic class, then passes of at most cfg.chunk numbers each
  std::vector<Piece> all;
  for (u32 i = 0; i < n; i++) split_classes(e, r[i].lo, r[i].hi, i, all);
  float sieve_ms = 0, test_ms = 0;
  const double chunk = (double)e->cfg.chunk;
  const u64 segspan = 30ull * e->cfg.seg_bytes;
  const u64 step = std::max<u64>(1, (u64)(chunk / (double)segspan)) * segspan;
  for (int cls = A_F64; cls <= A_FULL; cls++) {
    std::vector<Piece> batch;
    double in_batch = 0;
    for (const Piece& p0 : all) {
      if (p0.arith != cls) continue;
      u64 lo = p0.lo;
      for (;;) {
        const double rem = span_d(lo, p0.hi);
        const bool last = rem <= (double)step;
        const double take = last ? rem : (double)step;
        Piece part = p0;
        part.lo = lo;
        part.hi = last ? p0.hi : lo + step;
        if (in_batch + take > chunk && !batch.empty()) {
          if ((rc = run_pieces(e, batch, cls, to, sieve_ms, test_ms)) != WF_OK) return rc;
          batch.clear();
          in_batch = 0;
        }
        batch.push_back(part);
        in_batch += take;
        if (last) break;
        lo += step;
      }
    }
    if (!batch.empty() && (rc = run_pieces(e, batch, cls, to, sieve_ms, test_ms)) != WF_OK) return rc;
  }
  // results
  std::vector<DevAcc> acc(n);
  unsigned long long nev = 0;
  HIPCHK(hipMemcpyAsync(acc.data(), e->d_acc, n * sizeof(DevAcc), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(&nev, e->d_ctr + 1, sizeof nev, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  const u64 got = std::min<u64>(nev, e->dev_ev_cap);
  std::vector<DevEvent> dev(got);
  if (got) HIPCHK(hipMemcpy(dev.data(), e->d_ev, got * sizeof(DevEvent), hipMemcpyDeviceToHost));
  std::sort(dev.begin(), dev.end(), cmp_dev_event);
  for (u32 i = 0; i < n; i++) {
    wf_result& o = res[i];
    memset(&o, 0, sizeof o);
    o.primes = acc[i].primes;
    // the prime 2 is counted, never tested (wieferich.c:1142-1146)
    if (r[i].lo <= 2 && (r[i].hi == 0 || r[i].hi > 2)) o.primes++;
    o.tests = acc[i].tests;
    o.near = acc[i].near;
    o.flt = acc[i].flt;
    o.cks[0] = acc[i].cks;
    o.sumq[0] = acc[i].sumq;
    o.nev = acc[i].nev;
  }
  for (u64 k = 0; k < got && k < ev_cap; k++) {
    ev[k].type = dev[k].type;
    ev[k].j = 0;
    ev[k].p = dev[k].p;
    ev[k].q = dev[k].q;
    ev[k].r_hi = dev[k].r_hi;
    ev[k].r_lo = dev[k].r_lo;
  }
  e->last.sieve_ms = sieve_ms;
  e->last.test_ms = test_ms;
  e->last.total_ms = sieve_ms + test_ms;
  if (nev > e->dev_ev_cap) return WF_E_OVERFLOW;
  return WF_OK;
}

extern "C" int wf_scan(wf_engine* e, u64 lo, u64 hi, int64_t near_T, wf_result* out, wf_event* ev, u64 ev_cap,
                       int want_psp, u64* psp, u64 psp_cap) {
  if (!e || !out || (ev_cap && !ev)) return WF_E_ARG;
  if (want_psp) return WF_E_UNSUPPORTED;
  (void)psp;
  (void)psp_cap;
  wf_range r = {lo, hi};
  return scan_ranges(e, &r, 1, near_T, out, ev, ev_cap);
}

extern "C" int wf_scan_batch(wf_engine* e, const wf_range* r, u32 n, int64_t near_T, wf_result* per_range,
                             wf_event* ev, u64 ev_cap) {
  if (!e || !r || !per_range || (ev_cap && !ev)) return WF_E_ARG;
  return scan_ranges(e, r, n, near_T, per_range, ev, ev_cap);
}

extern "C" int wf_quotients(wf_engine* e, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n) {
  if (!e || !p || !q || !r_hi || !r_lo) return WF_E_ARG;
  if (n == 0) return WF_OK;
  for (u64 i = 0; i < n; i++)
    if (p[i] < 3 || !(p[i] & 1)) return WF_E_ARG;
  HIPCHK(hipSetDevice(e->dev));
  u64 *d = nullptr;
  if (hipMalloc(&d, 4 * n * sizeof(u64)) != hipSuccess) return WF_E_NOMEM;
  int rc = WF_OK;
  if (hipMemcpyAsync(d, p, n * sizeof(u64), hipMemcpyHostToDevice, e->st) != hipSuccess) rc = WF_E_DEVICE;
  if (rc == WF_OK) {
    hipLaunchKernelGGL(k_quot, dim3((u32)((n + 255) / 256)), dim3(256), 0, e->st, d, d + n, d + 2 * n, d + 3 * n, n);
    if (hipGetLastError() != hipSuccess) rc = WF_E_DEVICE;
  }
  if (rc == WF_OK && (hipMemcpyAsync(q, d + n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st) != hipSuccess ||
                      hipMemcpyAsync(r_hi, d + 2 * n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st) != hipSuccess ||
                      hipMemcpyAsync(r_lo, d + 3 * n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st) != hipSuccess ||
                      hipStreamSynchronize(e->st) != hipSuccess))
    rc = WF_E_DEVICE;
  hipFree(d);
  return rc;
}

extern "C" int wf_primes(wf_engine* e, u64 lo, u64 hi, u64* out, u64 cap, u64* count) {
  if (!e || !count || (cap && !out) || (hi != 0 && hi <= lo)) return WF_E_ARG;
  HIPCHK(hipSetDevice(e->dev));
  const u64 top = hi ? hi - 1 : 0xFFFFFFFFFFFFFFFFull;
  int rc;
  if ((rc = ensure_sieving_primes(e, std::max<u64>(isqrt_u64(top) + 1, 64))) != WF_OK) return rc;
  if ((rc = grow(e->d_acc, e->acc_cap, 1)) != WF_OK) return rc;
  HIPCHK(hipMemsetAsync(e->d_acc, 0, sizeof(DevAcc), e->st));
  std::vector<SegDesc> segs;
  build_segments(e, lo, hi, 0, segs);
  if ((rc = grow(e->d_segs, e->seg_cap, segs.size())) != WF_OK) return rc;
  const u64 bound = prime_bound(span_d(lo, hi));
  if ((rc = grow(e->d_primes, e->prime_cap, bound)) != WF_OK) return rc;
  if ((rc = grow(e->d_rid, e->rid_cap, e->prime_cap)) != WF_OK) return rc;
  HIPCHK(hipMemcpyAsync(e->d_segs, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemsetAsync(e->d_ctr, 0, sizeof(unsigned long long), e->st));
  const u32 nseg = (u32)segs.size();
  hipLaunchKernelGGL(k_sieve, dim3(std::min<u32>(nseg, (u32)e->prop.multiProcessorCount * 4)),
                     dim3(e->cfg.sieve_block), e->cfg.seg_bytes, e->st, e->d_segs, nseg, e->d_sp, e->d_spinv,
                     e->tierA, e->d_primes, e->d_rid, e->d_ctr, e->prime_cap, e->d_acc);
  HIPCHK(hipGetLastError());
  unsigned long long n = 0;
  HIPCHK(hipMemcpyAsync(&n, e->d_ctr, sizeof n, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  if (n > e->prime_cap) return WF_E_OVERFLOW;
  std::vector<u64> all(n);
  if (n) HIPCHK(hipMemcpy(all.data(), e->d_primes, n * sizeof(u64), hipMemcpyDeviceToHost));
  if (lo <= 2 && (hi == 0 || hi > 2)) all.push_back(2);
  std::sort(all.begin(), all.end());
  *count = all.size();
  for (u64 i = 0; i < all.size() && i < cap; i++) out[i] = all[i];
  return WF_OK;
}

extern "C" int wf_last_timing(wf_engine* e, wf_timing* t) {
  if (!e || !t) return WF_E_ARG;
  *t = e->last;
  return WF_OK;
}

extern "C" int wf_serve_start(wf_engine* e, u32 slots, u64 hbm_budget) {
  (void)e; (void)slots; (void)hbm_budget;
  return WF_E_UNSUPPORTED;
}
extern "C" int wf_serve_submit(wf_engine* e, u64 lo, u64 hi, u64* ticket) {
  (void)e; (void)lo; (void)hi; (void)ticket;
  return WF_E_UNSUPPORTED;
}
extern "C" int wf_serve_poll(wf_engine* e, u64 ticket, wf_result* out, int* done) {
  (void)e; (void)ticket; (void)out; (void)done;
  return WF_E_UNSUPPORTED;
}
extern "C" int wf_serve_stop(wf_engine* e) {
  (void)e;
  return WF_E_UNSUPPORTED;
}
