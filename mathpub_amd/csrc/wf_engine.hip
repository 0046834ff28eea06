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
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
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

__constant__ u32 c_R30[8] = {1, 7, 11, 13, 17, 19, 23, 29};
// c_M30[qi][o mod 30][c]: the m in [0, 30) with o + m q = R30[c] (mod 30), qi = the class
// of q mod 30 (an index into R30).
__constant__ unsigned char c_M30[8 * 30 * 8];
// c_QI[q mod 30] = index of q mod 30 in R30 (q coprime to 30)
__constant__ unsigned char c_QI[30];
// c_W30[r] = index of r in R30, -1 if r shares a factor with 30
__constant__ signed char c_W30[30] = {-1, 0, -1, -1, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1,
                                      -1, -1, 4, -1, 5, -1, -1, -1, 6, -1, -1, -1, -1, -1, 7};
// Wheel walk tables.  For a prime q of class qi and a multiple n = base + o = q m (base
// divisible by 30, t = o mod 30): c_WS[qi][t] = (j << 3) | k, where m + j is the first
// cofactor at or after m that is prime to 30 and R30[k] is its residue.  c_WCLS[qi] holds,
// in its 4-bit field k, the class (index in R30) of q R30[k] mod 30.
__constant__ unsigned char c_WS[8 * 30];
__constant__ u32 c_WCLS[8];
// gaps between consecutive residues of R30 (from R30[k] to R30[k + 1], 29 to 31), field k
#define WHEEL_GAPS 0x26424246u

// The wheel tables, copied to LDS by every workgroup that looks them up per lane
// (per-lane lookups in constant memory are vector memory loads with cache latency).
// m30 leads, 8-byte aligned, so that the eight m's of one (qi, o mod 30) are one read.
struct __attribute__((aligned(8))) WheelLds {
  unsigned char m30[8 * 30 * 8];
  unsigned char ws[8 * 30];
  u32 wcls[8];
  unsigned char qi[32];
  signed char w30[32];
  unsigned char r30[8];
  unsigned char bit_off[32];   // bit b of a bitmap word: 30 (b >> 3) + R30[b & 7], its offset from the word's first number
};

__device__ __forceinline__ void load_wheel(WheelLds& w) {
  for (u32 k = threadIdx.x; k < 8 * 30 * 8; k += blockDim.x) w.m30[k] = c_M30[k];
  for (u32 k = threadIdx.x; k < 8 * 30; k += blockDim.x) w.ws[k] = c_WS[k];
  if (threadIdx.x < 30) {
    w.qi[threadIdx.x] = c_QI[threadIdx.x];
    w.w30[threadIdx.x] = c_W30[threadIdx.x];
  }
  if (threadIdx.x < 8) {
    w.r30[threadIdx.x] = (unsigned char)c_R30[threadIdx.x];
    w.wcls[threadIdx.x] = c_WCLS[threadIdx.x];
  }
  if (threadIdx.x < 32) w.bit_off[threadIdx.x] = (unsigned char)(30u * (threadIdx.x >> 3) + c_R30[threadIdx.x & 7]);
}

// R30 packed, byte c = R30[c].
#define R30_PACKED 0x1D1713110D0B0701ull

// The eight m's of c_M30[qi][o mod 30] (byte c for class c) in one 64-bit LDS read, so that a
// prime's eight class starts wait on one lookup instead of eight.
__device__ __forceinline__ u64 class_ms(u64 o, int qi, const WheelLds& w) {
  u64 ms;
  __builtin_memcpy(&ms, &w.m30[(qi * 30 + (u32)o % 30u) * 8], sizeof ms);   // 8-byte aligned
  return ms;
}

// Byte index j (relative to the segment base) of the first mark of class c for a prime q
// whose first multiple at or after max(base, q^2) lies at offset o (numbers) from base;
// ms = class_ms(o, qi).  o < 2^32 at every call (o < q, or o < the span of a segment or
// fill run).  The numerator is a multiple of 30, so below 30 * 2^32 the division is exact:
// halve, then multiply by 15^-1 mod 2^32.  Larger starts lie beyond every segment and run.
__device__ __forceinline__ u64 class_start(u64 o, u32 q, u64 ms, int c) {
  const u32 o32 = (u32)o;
  const u32 m = (u32)(ms >> (8 * c)) & 0xFFu;
  const u32 r = (u32)(R30_PACKED >> (8 * c)) & 0xFFu;
  const u64 v = (u64)m * q + o32 - r;
  if (v >= (30ull << 32)) return ~0ull;   // beyond any segment or run (q close to 2^32)
  return (u64)((u32)(v >> 1) * 0xEEEEEEEFu);
}

__device__ __forceinline__ u32 rotr(u32 x, u32 r) { return (x >> r) | (x << ((32u - r) & 31u)); }   // r < 32

// Step of a wheel walk: gap * q, gap <= 6.  The 32-bit walk keeps q < 2^24 (one 24-bit
// multiply-add); the 64-bit walk takes any q.
__device__ __forceinline__ u32 wheel_step(u32 gap, u32 q) { return gap * (q & 0xFFFFFFu); }
__device__ __forceinline__ u64 wheel_step(u64 gap, u32 q) { return gap * q; }

// The multiples of a prime q > 5 on the wheel at offsets [o, span) from a base divisible by
// 30, in increasing order: f(byte, class) for each.  With base + o = q m, the multiple
// q (m + j) is on the wheel exactly when m + j is prime to 30, so the walk starts at the
// first such cofactor (c_WS) and then steps by q times the gaps between consecutive wheel
// residues; the class of each multiple rotates through c_WCLS[qi] in step.  Offsets are
// T: u32 needs q < 2^24 and span + 6 q < 2^32, u64 any q and span < 2^32.  With a few
// multiples per class this costs less than finding the eight class starts.
template <class T, class F>
__device__ __forceinline__ void wheel_walk(T o, u32 q, int qi, T span, const WheelLds& w, F f) {
  const u32 st = w.ws[qi * 30 + (u32)o % 30u];   // o < span < 2^32
  const u32 r = (st & 7u) * 4u;                   // field 0 = the first cofactor's residue
  u32 gaps = rotr(WHEEL_GAPS, r), cls = rotr(w.wcls[qi], r);
  T x = o + wheel_step((T)(st >> 3), q);
  while (x < span) {
    f((u32)x / 30u, (int)(cls & 7u));
    x += wheel_step((T)(gaps & 15u), q);
    gaps = rotr(gaps, 4);
    cls = rotr(cls, 4);
  }
}

// ---------------------------------------------------------------------------------
// sieve
// ---------------------------------------------------------------------------------
//
// A pass covers a list of segments.  Segment s holds the wheel-30 bitmap of
// [base, base + 30 nbytes) in LDS.  Sieving primes below the bucket threshold Q1 are
// applied directly per segment (tier A: q < 512, one wavefront per prime; tier B: one
// lane per prime).  Primes >= Q1 hit a segment at most a few times; k_bucket_fill walks
// each such prime over a run of consecutive segments once and leaves one 4-byte record
// (byte << 3 | class) per hit in the bucket of the segment it lands in.

#define RUN_MAX 1024      // segments per run (fill runs are cut shorter for smaller primes)
#define SEG_MAX 131072    // largest LDS segment in bytes (a bucket record is byte << 3 | class)
#define PRE_PERIOD 323323u  // 7 * 11 * 13 * 17 * 19: period of the presieve pattern in wheel bytes
#define PRE_COUNT 5u        // the presieve primes are the first five sieving primes
#ifndef FILL_STAGE
#define FILL_STAGE 16384  // records a fill workgroup stages in LDS (64 KiB)
#endif
#ifndef FILL_TARGET
#define FILL_TARGET (FILL_STAGE / 2)  // expected records per fill work item (the host sizes items to it)
#endif
#ifndef FILL_PER_PRIME
#define FILL_PER_PRIME 64 // records per prime and fill group the host sizes groups for
#endif
#ifndef FILL_PAIRS_DIV
#define FILL_PAIRS_DIV 1  // an item takes one lane per (prime, class) below blockDim / FILL_PAIRS_DIV primes
#endif
// Wave priority of the sieve-stage kernels.  In the pipelined path they share the CUs
// with the previous pass's test kernel, which alone keeps the VALU nearly saturated; a
// higher priority lets their (latency-bound) waves issue first, and the tests take the
// slots they leave.  Without a co-running kernel it changes nothing.
#ifndef SIEVE_PRIO
#define SIEVE_PRIO 3
#endif
#ifndef WALK_DIV
#define WALK_DIV 8        // fill: primes q > span / WALK_DIV walk their multiples instead of 8 class starts
#endif
#ifndef WHEEL_DIV
#define WHEEL_DIV 4       // sieve: primes q > seg_bytes / WHEEL_DIV walk their multiples
#endif
#ifndef FILL_BLOCK
#define FILL_BLOCK 1024   // threads per fill workgroup (16 wavefronts: two workgroups fill a CU)
#endif
#ifndef SIEVE_WAVES_EU
#define SIEVE_WAVES_EU 8  // sieve register budget: 8 wavefronts per SIMD (<= 64 VGPRs, no spills)
#endif

struct SegDesc {
  u64 base;       // multiple of 30
  u32 nbytes;     // wheel bytes (a multiple of 4, <= the LDS segment)
  u32 rid;        // range index
  u32 lo_off;     // valid numbers are base + [lo_off, hi_off)
  u32 hi_off;
  u32 nq;         // sieving primes used: q^2 <= the segment's last number (and q < bucket_min
                  // in prove mode)
  u32 prove;      // 1: sieved by the primes below bucket_min only; k_prove decides the survivors
};

// consecutive segments seg0 .. seg0 + nseg - 1 of one range piece, base = seg0's base
// (host side: the unit the fill groups are cut from)
struct RunDesc { u64 base; u32 seg0, nseg, nq, pad; };

// consecutive segments seg0 .. seg0 + nseg - 1 of one range piece (numbers from base),
// the segments lseg0 .. lseg0 + nseg - 1 of its fill group
struct FillRun { u64 base; u32 seg0, nseg, lseg0, pad; };

// one k_bucket_fill work item: sieving primes [i0, i1) over the fill runs fr0 .. fr0 + nfr - 1,
// nseg segments in all (<= RUN_MAX)
struct FillDesc { u32 fr0, nfr, i0, i1, nseg, fp; };   // fp: every run lies below 2^53

// A pass's fill plan as the host sends it, one array of u32 words: its fill groups (FillGroup,
// 8 words each, in work-item order), the group of every 64th work item, then the slice tables
// the groups share.  A group's primes [istart, bend) are cut at the ends words[slices],
// words[slices + 1], ... (one table serves every group of one prime range and group size);
// item0 is the index of its first work item.
struct FillGroup { u32 fr0, nfr, nseg, fp, istart, bend, slices, item0; };
static_assert(sizeof(FillGroup) == 8 * sizeof(u32), "a fill group is 8 words");

// Work item b of a pass: its group from the group of item b & ~63 and at most 63 steps on
// (almost always none), its slice from the group's table (the same item the host would have
// listed).
__device__ __forceinline__ FillDesc fill_item(const u32* __restrict__ fplan, u32 ngroups, u32 b) {
  const FillGroup* __restrict__ g = reinterpret_cast<const FillGroup*>(fplan);
  u32 lo = fplan[8 * ngroups + (b >> 6)];   // the last group whose first item is <= b
  while (lo + 1 < ngroups && g[lo + 1].item0 <= b) lo++;
  const FillGroup G = g[lo];
  const u32 k = b - G.item0;
  FillDesc it;
  it.fr0 = G.fr0;
  it.nfr = G.nfr;
  it.nseg = G.nseg;
  it.fp = G.fp;
  it.i0 = k ? fplan[G.slices + k - 1] : G.istart;
  it.i1 = min(fplan[G.slices + k], G.bend);
  return it;
}

struct DevAcc {     // per range; every field is a 64-bit sum (checksums are per range and base)
  unsigned long long primes, tests, near, flt, nev, npsp, nbpsw;
};

struct DevEvent { u32 type, rid, j, pad; u64 p, q, r_hi, r_lo; };

// The bases of the run: base j, its window w_j (the largest w <= 5 with b^(2^w - 1) < 2^32,
// as wieferich.c chooses it) and its powers b^d, d < 2^w.  Slot WF_MAXB holds the base of a
// wf_quotients_base call.
#define BPOW 32
__constant__ u32 c_base[WF_MAXB + 1];
__constant__ u32 c_bwin[WF_MAXB + 1];
__constant__ u32 c_bpow[(WF_MAXB + 1) * BPOW];
// Composite-base mode: the ladder primes l_s (ascending) with their powers l_s^d for the
// common window, and the exponent of l_s in base j (c_lexp[j * WF_MAXL + s]).
#define WF_MAXL 10
__constant__ u32 c_lbase[WF_MAXL];
__constant__ u32 c_lpow[WF_MAXL * BPOW];
__constant__ unsigned char c_lexp[WF_MAXB * WF_MAXL];

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

// The same for a base below 2^53 (exact as a double), from rq = 1/q rounded to nearest:
// t = floor(base * rq) is within one of floor(base / q) (the relative errors of rq and of the
// product are below 2^-52 each, and base / q < 2^53 / q), so the remainder base - t q, which
// one FMA gives exactly (an integer below 2q in magnitude), needs one correction each way.
// chk_sq (uniform): q^2 may lie at or above the base.
__device__ __forceinline__ bool first_offset_fp(double bd, u64 base, u32 q, double rq, u64 span, bool chk_sq,
                                                u64& o) {
  if (chk_sq) {
    const u64 qq = (u64)q * q;
    if (qq >= base) {
      o = qq - base;
      return o < span;
    }
  }
  const double qd = (double)q;
  const double t = __builtin_floor(bd * rq);
  double r = __builtin_fma(-t, qd, bd);
  r = r < 0.0 ? r + qd : r;
  r = r >= qd ? r - qd : r;
  const u32 ri = (u32)r;
  o = ri ? q - ri : 0;
  return o < span;
}

// Every hit of an item's primes in its fill group: hit(segment of the group, byte offset
// within the run, class), or hit4 for four hits of one class at once (their LDS atomics are
// then in flight together instead of each waiting for the last).  A lane takes one prime
// (all 8 classes) or, when the slice is short, one (prime, class) pair; primes with few hits
// per run walk their multiples.
template <class F, class F4>
__device__ __forceinline__ void fill_walk(const FillDesc& it, const FillRun* __restrict__ fruns,
                                          const u32* __restrict__ sp, const u64* __restrict__ spinv,
                                          const double* __restrict__ sprcp, const WheelLds& wl, int sbshift,
                                          bool pairs, F hit, F4 hit4) {
  const u32 nwork = pairs ? (it.i1 - it.i0) * 8 : it.i1 - it.i0;
  // Placement: for an item below 2^53 (its primes are below 2^26.5) by neg_mod_fp from the
  // FP64 reciprocals, else by Barrett reduction.  The next work unit's prime and constant are
  // loaded while this one is walked.
  const bool fpf = it.fp != 0;
  const u64* __restrict__ mtab = fpf ? reinterpret_cast<const u64*>(sprcp) : spinv;
  u32 w = threadIdx.x;
  u32 qn = 0;
  u64 mn = 0;
  if (w < nwork) {
    const u32 i = it.i0 + (pairs ? w >> 3 : w);
    qn = sp[i];
    mn = mtab[i];
  }
  for (; w < nwork; w += blockDim.x) {
    const u32 q = qn;
    const u64 minv = mn;
    if (w + blockDim.x < nwork) {
      const u32 i = it.i0 + (pairs ? (w + blockDim.x) >> 3 : w + blockDim.x);
      qn = sp[i];
      mn = mtab[i];
    }
    const int qi = wl.qi[q % 30];
    const int c0 = pairs ? (int)(w & 7) : 0, c1 = pairs ? c0 + 1 : 8;
    for (u32 r = 0; r < it.nfr; r++) {
      const FillRun fr = fruns[it.fr0 + r];
      const u64 nbytes = (u64)fr.nseg << sbshift;
      const u64 span = 30ull * nbytes;
      u64 o;
      if (fpf) {
        const u64 qq = (u64)q * q;
        o = qq >= fr.base ? qq - fr.base : neg_mod_fp((double)fr.base, q, __builtin_bit_cast(double, minv));
        if (o >= span) continue;
      } else if (!first_offset(fr.base, q, minv, span, o)) {
        continue;
      }
      if (!pairs && (u64)WALK_DIV * q > span) {
        wheel_walk<u64>(o, q, qi, span, wl, [&](u32 j, int c) { hit(fr.lseg0 + (j >> sbshift), j, (u32)c); });
        continue;
      }
      const u64 ms = class_ms(o, qi, wl);
      for (int c = c0; c < c1; c++) {
        u64 j = class_start(o, q, ms, c);
        for (; j + 3ull * q < nbytes; j += 4ull * q) {
          u32 sg[4], jj[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            jj[k] = (u32)(j + (u64)k * q);
            sg[k] = fr.lseg0 + (jj[k] >> sbshift);
          }
          hit4(sg, jj, (u32)c);
        }
        for (; j < nbytes; j += q) hit(fr.lseg0 + (u32)(j >> sbshift), (u32)j, (u32)c);
      }
    }
  }
}

// Large primes -> bucket records.  A work item covers a slice of the large primes over one
// fill group: consecutive chunks of segments (of one range, or of several short ranges),
// RUN_MAX segments at most, sized by the host to about FILL_TARGET expected records.  A
// prime's table entry is read once per group, and one Barrett reduction per run places it.
// One walk over the hits: each segment of the group owns C = FILL_STAGE / nseg (a power of
// two, about twice the expected count) stage slots in LDS, taken by an LDS atomic; a record
// beyond them (rare) reserves its bucket slot with a global atomic and is written at once.
// Then one global atomic per segment reserves bucket room for its staged records (and clamps
// the count to the room left), and the workgroup copies them out with consecutive lanes on
// consecutive addresses.  One workgroup per item: a grid of resident workgroups that loop
// over the items measured 12% slower on the sieve stage.
__global__ void __launch_bounds__(FILL_BLOCK) k_bucket_fill(const u32* __restrict__ fplan, u32 ngroups,
                                                     const FillRun* __restrict__ fruns,
                                                     const u32* __restrict__ sp, const u64* __restrict__ spinv,
                                                     const double* __restrict__ sprcp, u32* __restrict__ bcnt, u32* __restrict__ bdata,
                                                     u32 bcap, int sbshift, u32* __restrict__ flags) {
  __shared__ u32 cur[RUN_MAX];    // records per segment (stage slots taken), then the count to copy
  __shared__ union {
    u32 gseg[RUN_MAX];            // during the walk: group segment -> segment of the pass
    u64 dst[RUN_MAX];             // for the copy: bucket index of the segment's first staged record
  } sw;
  __shared__ u32 stage[FILL_STAGE];
  __shared__ WheelLds wl;
  if (SIEVE_PRIO) __builtin_amdgcn_s_setprio(SIEVE_PRIO);
  load_wheel(wl);
  const FillDesc it = fill_item(fplan, ngroups, blockIdx.x);
  const u32 smask = (1u << sbshift) - 1;
  const int cshift = 31 - __builtin_clz(FILL_STAGE / it.nseg);   // it.nseg <= RUN_MAX < FILL_STAGE
  const u32 C = 1u << cshift;
  for (u32 k = threadIdx.x; k < it.nseg; k += blockDim.x) cur[k] = 0;
  for (u32 r = 0; r < it.nfr; r++) {
    const FillRun fr = fruns[it.fr0 + r];
    for (u32 k = threadIdx.x; k < fr.nseg; k += blockDim.x) sw.gseg[fr.lseg0 + k] = fr.seg0 + k;
  }
  __syncthreads();
  // a record to its stage slot, or past the segment's room straight to its bucket
  auto place = [&](u32 sg, u32 slot, u32 rec) {
    if (slot < C) {
      stage[(sg << cshift) + slot] = rec;
    } else {
      const u32 g = sw.gseg[sg];
      const u32 pos = atomicAdd(&bcnt[g], 1u);
      if (pos < bcap) bdata[(u64)g * bcap + pos] = rec;
      else atomicOr(flags, 1u);
    }
  };
  // few primes with many hits each: one lane per (prime, class); else one lane per prime
  const bool pairs = (it.i1 - it.i0) * FILL_PAIRS_DIV < blockDim.x;
  fill_walk(it, fruns, sp, spinv, sprcp, wl, sbshift, pairs, [&](u32 sg, u32 j, u32 c) {
    place(sg, atomicAdd(&cur[sg], 1u), ((j & smask) << 3) | c);
  }, [&](const u32 (&sg)[4], const u32 (&jj)[4], u32 c) {
    u32 slot[4];
#pragma unroll
    for (int k = 0; k < 4; k++) slot[k] = atomicAdd(&cur[sg[k]], 1u);
#pragma unroll
    for (int k = 0; k < 4; k++) place(sg[k], slot[k], ((jj[k] & smask) << 3) | c);
  });
  __syncthreads();
  // one bucket reservation per segment for its staged records (it.nseg <= RUN_MAX = the
  // block size: one segment per thread); what does not fit is flagged and left out
  static_assert(RUN_MAX <= FILL_BLOCK, "one segment per thread");
  const u32 k = threadIdx.x;
  u32 n = 0;
  u64 b = 0;
  if (k < it.nseg) {
    const u32 g = sw.gseg[k];
    n = min(cur[k], C);
    u32 pos = 0;
    if (n) {
      pos = atomicAdd(&bcnt[g], n);
      if (pos + n > bcap) {
        atomicOr(flags, 1u);
        n = pos < bcap ? bcap - pos : 0;
      }
    }
    b = (u64)g * bcap + pos;
  }
  __syncthreads();   // every gseg entry is read before dst overwrites it
  if (k < it.nseg) {
    sw.dst[k] = b;
    cur[k] = n;
  }
  __syncthreads();
  if (C >= 128) {
    // 64 or more expected records per segment: a wavefront per segment copies its filled
    // slots only, consecutive lanes on consecutive addresses
    const u32 lane = threadIdx.x & 63;
    for (u32 s = threadIdx.x >> 6; s < it.nseg; s += blockDim.x >> 6) {
      const u32 n = cur[s];
      const u32* src = stage + (s << cshift);
      u32* dst = bdata + sw.dst[s];
      for (u32 i = lane; i < n; i += 64) dst[i] = src[i];
    }
  } else {
    // fewer: the threads cover the slots (C <= 64 divides the block), each keeping its slot
    // index i while it steps over the segments
    const u32 i = threadIdx.x & (C - 1), step = blockDim.x >> cshift;
    for (u32 s = threadIdx.x >> cshift; s < it.nseg; s += step)
      if (i < cur[s]) bdata[sw.dst[s] + i] = stage[(s << cshift) + i];
  }
}

// Valid-bit mask of LDS word w (bytes 4w .. 4w+3) for numbers base + [lo_off, hi_off).
__device__ __forceinline__ u32 word_mask(u32 w, u32 lo_off, u32 hi_off) {
  const u32 w0 = 120u * w;
  if (w0 >= lo_off && w0 + 120u <= hi_off) return 0xFFFFFFFFu;
  u32 m = 0;   // boundary words only: a rolled loop keeps the register count down
#pragma unroll 1
  for (int b = 0; b < 32; b++) {
    const u32 off = w0 + 30u * (b >> 3) + c_R30[b & 7];
    if (off >= lo_off && off < hi_off) m |= 1u << b;
  }
  return m;
}

// Sieve one segment per workgroup iteration and write its primes, in ascending order,
// as offsets from the segment base into out[s * pcap ...]; cnt[s] = how many.
__attribute__((amdgpu_waves_per_eu(SIEVE_WAVES_EU)))
__global__ void __launch_bounds__(1024) k_sieve(const SegDesc* __restrict__ segs, u32 nseg,
                                                const u32* __restrict__ sp, const u64* __restrict__ spinv,
                                                const double* __restrict__ sprcp, u32 tierA, u32 q1idx,
                                                const u32* __restrict__ bcnt, const u32* __restrict__ bdata,
                                                u32 bcap,
                                                u32* __restrict__ out, u32* __restrict__ cnt, u32 pcap,
                                                DevAcc* __restrict__ acc, u32* __restrict__ flags,
                                                u32* __restrict__ bitmap, u32 bmw,
                                                const u32* __restrict__ pre, u32 pre_words) {
  extern __shared__ u32 seg[];
  __shared__ u32 wtot[16];
  __shared__ WheelLds wl;
  if (SIEVE_PRIO) __builtin_amdgcn_s_setprio(SIEVE_PRIO);
  load_wheel(wl);   // the first barrier below orders it before every lookup
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  for (u32 s = blockIdx.x; s < nseg; s += gridDim.x) {
    const SegDesc d = segs[s];
    const u32 nwords = d.nbytes >> 2;
    const u64 span = 30ull * d.nbytes;
    // start from the pattern of the presieve primes 7, 11, 13, 17, 19 (period 323323 bytes)
    {
      // four loads in flight per lane before the LDS writes
      const u32 ph = (u32)((d.base / 30) % PRE_PERIOD);
      const u32* pw = pre + (u64)(ph & 3) * pre_words + (ph >> 2);
      const u32 bd = blockDim.x;
      for (u32 w = threadIdx.x; w < nwords; w += 4 * bd) {
        u32 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = w + k * bd < nwords ? pw[w + k * bd] : 0u;
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (w + k * bd < nwords) seg[w + k * bd] = v[k];
      }
    }
    // 7, 11, 13, 17, 19 themselves are prime; thread 0 stored word 0 itself just above
    if (d.base == 0 && threadIdx.x == 0) seg[0] &= ~0x3Eu;
    __syncthreads();
    const u32 nsmall = min(d.nq, q1idx);
    // Placement (uniform per segment): below 2^53 from the FP64 reciprocal table, else by
    // Barrett reduction.  The q^2 test is needed only if the largest prime applied reaches it.
    const bool fp = d.base < (1ull << 53);
    const double bd = (double)d.base;
    const u64 qtop = nsmall ? sp[nsmall - 1] : 0;
    const bool chk_sq = qtop * qtop >= d.base;
    const u64* __restrict__ mtab = fp ? reinterpret_cast<const u64*>(sprcp) : spinv;
    auto place = [&](u32 q, u64 m, u64& o) {
      return fp ? first_offset_fp(bd, d.base, q, __builtin_bit_cast(double, m), span, chk_sq, o)
                : first_offset(d.base, q, m, span, o);
    };
    // tier B (q >= 512): below 2^53 the one-sided FP64 placement (neg_mod_fp, q < 2^31)
    const bool fpb = fp && qtop < 0x80000000ull;
    auto place_b = [&](u32 q, u64 m, u64& o) {
      if (!fpb) return place(q, m, o);
      if (chk_sq) {
        const u64 qq = (u64)q * q;
        if (qq >= d.base) {
          o = qq - d.base;
          return o < span;
        }
      }
      o = neg_mod_fp(bd, q, __builtin_bit_cast(double, m));
      return o < span;
    };
    // tier A: one wavefront per prime, lanes stride over the marks of each class.  A lane's
    // step, 64 q bytes, is a multiple of 4: its marks of one class keep their byte within the
    // word, so the bit is set up once and the loop advances a word index.
    const u32 na = min(tierA, nsmall);
    u32 qa = PRE_COUNT + wave < na ? sp[PRE_COUNT + wave] : 0;
    u64 ma = PRE_COUNT + wave < na ? mtab[PRE_COUNT + wave] : 0;
    for (u32 i = PRE_COUNT + wave; i < na; i += nwaves) {
      const u32 q = qa;
      const u64 m = ma;
      if (i + nwaves < na) {      // the wavefront's next prime, loaded while this one is sieved
        qa = sp[i + nwaves];
        ma = mtab[i + nwaves];
      }
      u64 o;
      if (!place(q, m, o)) continue;
      const int qi = wl.qi[q % 30];
      const u32 jl = (u32)class_start(o, q, class_ms(o, qi, wl), lane & 7);   // lane c (< 8): class c's start
      for (int c = 0; c < 8; c++) {
        const u32 j = __shfl(jl, c, 64) + (u32)lane * q;   // q < 512: starts lie within the span
        if (j >= d.nbytes) continue;
        const u32 bit = 1u << (((j & 3) << 3) | c);
        for (u32 w = j >> 2; w < nwords; w += 16u * q) atomicOr(&seg[w], bit);
      }
    }
    // tier B: one lane per prime; the next prime's q and Barrett constant are loaded
    // while this one is sieved
    {
      u32 i = tierA + threadIdx.x;
      u32 qn = i < nsmall ? sp[i] : 0;
      u64 mn = i < nsmall ? mtab[i] : 0;
      for (; i < nsmall; i += blockDim.x) {
        const u32 q = qn;
        const u64 m = mn;
        if (i + blockDim.x < nsmall) {
          qn = sp[i + blockDim.x];
          mn = mtab[i + blockDim.x];
        }
        u64 o;
        if (!place_b(q, m, o)) continue;
        const int qi = wl.qi[q % 30];
        if ((u64)WHEEL_DIV * q > d.nbytes) {     // fewer than WHEEL_DIV marks per class
          auto mark = [&](u32 j, int c) { atomicOr(&seg[j >> 2], 1u << (((j & 3) << 3) | c)); };
          if (q < (1u << 24)) wheel_walk<u32>((u32)o, q, qi, (u32)span, wl, mark);
          else wheel_walk<u64>(o, q, qi, span, wl, mark);
          continue;
        }
        const u64 ms = class_ms(o, qi, wl);
        if (8u * q <= d.nbytes) {
          // eight or more marks per class: the marks j0 + (4k + t) q of one t keep their byte
          // within the word (4 q bytes = q words), so each of the four runs sets one bit
          // through a word index
#pragma unroll 1
          for (int c = 0; c < 8; c++) {
            const u32 j0 = (u32)class_start(o, q, ms, c);
            if (j0 >= d.nbytes) continue;
#pragma unroll
            for (u32 t = 0; t < 4; t++) {
              const u32 j = j0 + t * q;
              const u32 bit = 1u << (((j & 3) << 3) | c);
              for (u32 w = j >> 2; w < nwords; w += q) atomicOr(&seg[w], bit);
            }
          }
          continue;
        }
#pragma unroll 1
        for (int c = 0; c < 8; c++) {
          for (u32 j = (u32)class_start(o, q, ms, c); j < d.nbytes; j += q)
            atomicOr(&seg[j >> 2], 1u << (((j & 3) << 3) | c));
        }
      }
    }
    // large primes: this segment's bucket, four records per 16-byte load (bcap is a multiple
    // of 64), the next load issued before the current records are applied
    if (d.nq > q1idx) {
      const u32 nrec = min(bcnt[s], bcap);
      const u32* rec = bdata + (u64)s * bcap;
      const uint4* r4 = reinterpret_cast<const uint4*>(rec);
      const u32 n4 = nrec >> 2;
      auto apply = [&](u32 v) { atomicOr(&seg[v >> 5], 1u << ((((v >> 3) & 3) << 3) | (v & 7))); };
      u32 k = threadIdx.x;
      uint4 nx = k < n4 ? r4[k] : make_uint4(0, 0, 0, 0);
      for (; k < n4; k += blockDim.x) {
        const uint4 v = nx;
        if (k + blockDim.x < n4) nx = r4[k + blockDim.x];
        apply(v.x);
        apply(v.y);
        apply(v.z);
        apply(v.w);
      }
      for (u32 t = (n4 << 2) + threadIdx.x; t < nrec; t += blockDim.x) apply(rec[t]);
    }
    __syncthreads();
    // composite bitmap for the strong-pseudoprime pass
    if (bitmap)
      for (u32 w = threadIdx.x; w < nwords; w += blockDim.x) bitmap[(u64)s * bmw + w] = seg[w];
    // extraction: wave w owns words [w M, (w + 1) M); lanes read consecutive words
    const u32 M = (nwords + nwaves - 1) / nwaves;
    const u32 wb = min((u32)wave * M, nwords), we = min(wb + M, nwords);
    u32 c0 = 0;
    for (u32 w = wb + lane; w < we; w += 64) c0 += __popc(~seg[w] & word_mask(w, d.lo_off, d.hi_off));
    // the primes 3 and 5 are not on the wheel; they lead segment 0
    u32 extra = 0;
    if (d.base == 0) {
      if (3 >= d.lo_off && 3 < d.hi_off) extra |= 1;
      if (5 >= d.lo_off && 5 < d.hi_off) extra |= 2;
    }
    const u32 wsum = (u32)wave_sum((u64)c0) + (wave == 0 ? __popc(extra) : 0);
    if (lane == 0) wtot[wave] = wsum;
    __syncthreads();
    u32 pos = 0, total = 0;
    for (int k = 0; k < nwaves; k++) {
      const u32 t = wtot[k];
      if (k < wave) pos += t;
      total += t;
    }
    u32* o = out + (u64)s * pcap;
    if (wave == 0) {
      if (lane == 0) {
        u32 k = 0;
        if ((extra & 1) && k < pcap) o[k++] = 3;
        if ((extra & 2) && k < pcap) o[k++] = 5;
      }
      pos += __popc(extra);
    }
    // a segment whose primes all fit its region (always, unless the flag below is raised)
    // stores without a bound check
    const bool fits = total <= pcap;
    for (u32 w0 = wb; w0 < we; w0 += 64) {
      const u32 w = w0 + lane;
      u32 v = w < we ? (~seg[w] & word_mask(w, d.lo_off, d.hi_off)) : 0;
      const u32 c = __popc(v);
      const u32 incl = wave_incl_scan(c);
      u32 k = pos + incl - c;
      const u32 w120 = 120u * w;
      if (fits) {
        for (; v; v &= v - 1) o[k++] = w120 + wl.bit_off[__builtin_ctz(v)];
      } else {
        for (; v; v &= v - 1, k++)
          if (k < pcap) o[k] = w120 + wl.bit_off[__builtin_ctz(v)];
      }
      pos += __shfl(incl, 63, 64);
    }
    if (threadIdx.x == 0) {
      cnt[s] = min(total, pcap);
      if (total > pcap) atomicOr(flags, 2u);
      if (total && !d.prove) atomicAdd(&acc[d.rid].primes, (unsigned long long)total);
    }
    __syncthreads();
  }
}

// Ordered compaction over a 256-thread workgroup: the values x of the lanes with keep set go
// to dst[0..total) in thread order.  Every thread calls it; wtot holds the 4 wavefront counts.
__device__ __forceinline__ u32 block_compact(bool keep, u32 x, u32* __restrict__ dst, u32* wtot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const u64 m = __ballot(keep);
  if (lane == 0) wtot[wave] = (u32)__popcll(m);
  __syncthreads();
  u32 before = 0, total = 0;
  for (int w = 0; w < 4; w++) {
    const u32 c = wtot[w];
    if (w < wave) before += c;
    total += c;
  }
  if (keep) dst[before + (u32)__popcll(m & ((1ull << lane) - 1))] = x;
  __syncthreads();   // wtot and dst are read again by the caller's next step
  return total;
}

// Prove mode.  The segment was sieved by the primes below bucket_min only, so its list holds
// every prime and the composites whose prime factors all exceed that bound.  A survivor below
// q_next^2 (q_next = the first sieving prime left out) is prime; any other is decided by
// BPSW, which has no pseudoprime below 2^64.
//
// k_prove: a work item is one tile of PROVE_TILE survivors of one segment (items run on
// every CU, whatever the number of segments).  Phase 1 runs the base-2 strong test and keeps
// the passes in LDS, in order; phase 2 runs the strong Lucas test on them with every lane
// busy and writes the primes to the front of the tile, in order; tcnt holds the tile's count.
// k_prove_gather then closes the gaps: per segment, tile after tile, it moves each tile's
// primes down to the running total (each move stays below the next tile's start).
#define PROVE_TILE 1024

__global__ void __launch_bounds__(256) k_prove(const SegDesc* __restrict__ segs, u32 nseg, u32 ptps,
                                               const u32* __restrict__ sp, u32* __restrict__ off,
                                               const u32* __restrict__ cnt, u32 pcap, u32* __restrict__ tcnt) {
  __shared__ u32 queue[PROVE_TILE];
  __shared__ u32 wtot[4];
  const u64 nitems = (u64)nseg * ptps;
  for (u64 t = blockIdx.x; t < nitems; t += gridDim.x) {
    const u32 s = (u32)(t / ptps);
    const u32 t0 = (u32)(t - (u64)s * ptps) * PROVE_TILE;
    const SegDesc d = segs[s];
    if (!d.prove) continue;                         // uniform across the workgroup
    const u32 nin = cnt[s];
    if (t0 >= nin) continue;
    const u32 t1 = min(nin, t0 + PROVE_TILE);
    u32* o = off + (u64)s * pcap;
    const u64 qn = sp[d.nq];
    const u64 lim = qn * qn;                        // q_next < 2^32
    u32 nq = 0;
    for (u32 j0 = t0; j0 < t1; j0 += 256) {
      const u32 i = j0 + threadIdx.x;
      u32 x = 0;
      bool keep = false;
      if (i < t1) {
        x = o[i];
        const u64 n = d.base + x;
        keep = n < lim || sprp2_64(n);
      }
      nq += block_compact(keep, x, queue + nq, wtot);
    }
    u32 wp = 0;
    for (u32 k0 = 0; k0 < nq; k0 += 256) {
      const u32 k = k0 + threadIdx.x;
      u32 x = 0;
      bool keep = false;
      if (k < nq) {
        x = queue[k];
        const u64 n = d.base + x;
        keep = n < lim || slprp_64(n);
      }
      wp += block_compact(keep, x, o + t0 + wp, wtot);
    }
    if (threadIdx.x == 0) tcnt[t] = wp;
    __syncthreads();   // the queue is refilled by the next item
  }
}

__global__ void __launch_bounds__(PROVE_TILE) k_prove_gather(const SegDesc* __restrict__ segs, u32 nseg, u32 ptps,
                                                             u32* __restrict__ off, u32* __restrict__ cnt, u32 pcap,
                                                             const u32* __restrict__ tcnt, DevAcc* __restrict__ acc,
                                                             unsigned long long* __restrict__ proved) {
  for (u32 s = blockIdx.x; s < nseg; s += gridDim.x) {
    const SegDesc d = segs[s];
    if (!d.prove) continue;
    const u32 ntile = (cnt[s] + PROVE_TILE - 1) / PROVE_TILE;
    u32* o = off + (u64)s * pcap;
    u32 pos = 0;
    for (u32 k = 0; k < ntile; k++) {
      const u32 c = tcnt[(u64)s * ptps + k];
      const u32 src = k * PROVE_TILE;
      if (src != pos) {                             // uniform
        const u32 x = threadIdx.x < c ? o[src + threadIdx.x] : 0;
        __syncthreads();                            // every read before any write of this move
        if (threadIdx.x < c) o[pos + threadIdx.x] = x;
        __syncthreads();
      }
      pos += c;
    }
    if (threadIdx.x == 0) {
      cnt[s] = pos;
      if (pos) {
        atomicAdd(&acc[d.rid].primes, (unsigned long long)pos);
        atomicAdd(proved, (unsigned long long)pos);
      }
    }
    __syncthreads();   // cnt[s] is read above by every thread of the workgroup
  }
}

// ---------------------------------------------------------------------------------
// tests
// ---------------------------------------------------------------------------------

#ifndef TEST_TILE
#define TEST_TILE 1024   // primes per test work item (4 rounds of 256 lanes)
#endif

// arithmetic classes: FP64 p-adic (p < F64_PMAX = 9 * 2^47), 32-bit Montgomery with 2 / 3 / 4 limbs,
// 4 limbs keeping the carry bit (p >= 2^63), radix-2^28 Montgomery (p >= 2^47)
enum Arith { A_F64 = 0, A_M2, A_M3, A_M4, A_FULL, A_M28, A_NCLS };

struct TestOut {
  DevAcc* acc;               // per range
  unsigned long long* cks;   // per range and base: [rid * nb + j]
  unsigned long long* sumq;
  DevEvent* ev;
  u64 ev_cap;
  unsigned long long* nev;   // total events (device counter)
  i64 near_T;                // near-hit bound of every range, unless nearv is set
  const i64* nearv;          // per range (serving: per slot), or nullptr
  u64 fault_p;               // testing only: the residue of this p is taken as r + 1 (0 = off)
  u32 nb;                    // bases of the run
};

__device__ __forceinline__ void push_event(const TestOut& o, u32 type, u32 rid, u32 j, u64 p, u64 q, u64 rhi,
                                           u64 rlo) {
  const unsigned long long k = atomicAdd(o.nev, 1ull);
  atomicAdd(&o.acc[rid].nev, 1ull);
  if (k < o.ev_cap) {
    DevEvent e;
    e.type = type; e.rid = rid; e.j = j; e.pad = 0; e.p = p; e.q = q; e.r_hi = rhi; e.r_lo = rlo;
    o.ev[k] = e;
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

// |A| <= T for the signed quotient A = q (q <= p/2) or q - p (q > p/2), wieferich.c's
// near-hit test: |A| = min(q, p - q) for every p and q < p.
__device__ __forceinline__ bool near_hit(u64 p, u64 q, i64 T) {
  const u64 A = q < p - q ? q : p - q;
  return A <= (u64)T;
}

// Account one tested (prime, base) pair exactly as finish_lanes (wieferich.c:1051-1075)
// and add the lane's terms to its running sums.  r = r_hi:r_lo, q = (r - 1) p^-1 mod 2^64.
struct LaneSums { u32 tests, flt, near; u64 cks, sumq; };   // counts per lane and launch fit 32 bits

__device__ __forceinline__ void account(const TestOut& o, u32 rid, u32 j, u64 p, u64 rhi, u64 rlo, LaneSums& s,
                                        i64 nT) {
  if (p == o.fault_p) {      // injected fault: r + 1 is not 1 mod p, so the FLT check must catch it
    rlo += 1;
    rhi += rlo == 0 ? 1 : 0;
  }
  const u64 pinv = inv64(p);
  const u64 q = (rlo - 1) * pinv;
  s.tests++;
  if (q >= p) {
    s.flt++;
    push_event(o, WF_EV_FLT, rid, j, p, q, rhi, rlo);
    return;
  }
  s.cks += d_mix64(q ^ (p * GOLD));
  s.sumq += q;
  if (q == 0) {
    push_event(o, WF_EV_FOUND, rid, j, p, q, rhi, rlo);
  } else if (nT > 0 && near_hit(p, q, nT)) {
    s.near++;
    push_event(o, WF_EV_NEAR, rid, j, p, q, rhi, rlo);
  }
}

// The same from the ladder's p-adic digits of r = a + b p (pd = p as a double): for prime p
// not dividing the base, r = 1 (mod p), so a = 1 and q = b mod p (f64_quotient); p^-1 and r
// are only formed for an event or a failed FLT check.
__device__ __forceinline__ void account_padic(const TestOut& o, u32 rid, u32 j, u64 p, double pd, double a,
                                              double b, LaneSums& s, i64 nT) {
  u64 q;
  if (p != o.fault_p && f64_quotient(a, b, pd, q)) {   // q < p: the FLT check holds
    s.tests++;
    s.cks += d_mix64(q ^ (p * GOLD));
    s.sumq += q;
    if (q == 0) {
      push_event(o, WF_EV_FOUND, rid, j, p, 0, 0, 1);
    } else if (nT > 0 && near_hit(p, q, nT)) {
      s.near++;
      const u64 lo = q * p, hi = __umul64hi(q, p);
      const u64 rlo = lo + 1;
      push_event(o, WF_EV_NEAR, rid, j, p, q, hi + (rlo < lo), rlo);
    }
    return;
  }
  F64Mod m;
  m.p = pd;
  m.ip = 1.0 / pd;
  f64_normalise(a, b, m);
  u64 rhi, rlo;
  padic_to_r(p, a, b, rhi, rlo);
  account(o, rid, j, p, rhi, rlo, s, nT);
}

// Wavefront sums of a (range, base)'s lane totals -> one atomic per counter (lane 0).
__device__ __forceinline__ void flush_sums(const TestOut& o, u32 rid, u32 j, LaneSums& s) {
  const u64 t = wave_sum((u64)s.tests), f = wave_sum((u64)s.flt), n = wave_sum((u64)s.near);
  const u64 c = wave_sum(s.cks), q = wave_sum(s.sumq);
  if ((threadIdx.x & 63) == 0 && t) {
    DevAcc* a = &o.acc[rid];
    atomicAdd(&a->tests, t);
    if (f) atomicAdd(&a->flt, f);
    if (n) atomicAdd(&a->near, n);
    atomicAdd(&o.cks[(u64)rid * o.nb + j], c);
    atomicAdd(&o.sumq[(u64)rid * o.nb + j], q);
  }
  s.tests = s.flt = s.near = s.cks = s.sumq = 0;
}

// true if p divides the base b (then b^(p-1) = 0 mod p and the pair is skipped, as
// process_batch does, wieferich.c:1097)
__device__ __forceinline__ bool divides_base(u64 p, u32 b) { return p <= b && b % (u32)p == 0; }

// r = b^(p-1) mod p^2 for one odd p of the class: base 2 by its own ladders, any other base
// by windows of w bits with the powers pw[d] = b^d.
template <int ARITH>
__device__ __forceinline__ void residue(u64 p, int eb, u32 b, int w, const u32* pw, u64& rhi, u64& rlo) {
  const u64 e = p - 1;
  if (ARITH == A_F64) {
    const double pd = (double)p;
    double a0, b0, a1, b1;
    if (b == 2) {
      fp64_pow2x2<5>(pd, e, 3.0, 2, eb, a0, b0, a1, b1);
    } else {
      switch (w) {
        case 5: fp64_powc_x2<5>(pd, e, 3.0, 2, eb, pw, pw, a0, b0, a1, b1); break;
        case 4: fp64_powc_x2<4>(pd, e, 3.0, 2, eb, pw, pw, a0, b0, a1, b1); break;
        case 3: fp64_powc_x2<3>(pd, e, 3.0, 2, eb, pw, pw, a0, b0, a1, b1); break;
        case 2: fp64_powc_x2<2>(pd, e, 3.0, 2, eb, pw, pw, a0, b0, a1, b1); break;
        default: fp64_powc_x2<1>(pd, e, 3.0, 2, eb, pw, pw, a0, b0, a1, b1); break;
      }
    }
    F64Mod m;
    m.p = pd;
    m.ip = 1.0 / pd;
    f64_normalise(a0, b0, m);
    padic_to_r(p, a0, b0, rhi, rlo);
  } else if (ARITH == A_M28) {
    u32 x[4];
    if (b == 2) {
      m28_pow2(p, e, eb, x);
    } else {
      switch (w) {
        case 5: m28_powc<5>(p, e, eb, pw, x); break;
        case 4: m28_powc<4>(p, e, eb, pw, x); break;
        case 3: m28_powc<3>(p, e, eb, pw, x); break;
        case 2: m28_powc<2>(p, e, eb, pw, x); break;
        default: m28_powc<1>(p, e, eb, pw, x); break;
      }
    }
    rlo = (u64)x[0] | ((u64)x[1] << 32);
    rhi = (u64)x[2] | ((u64)x[3] << 32);
  } else {
    // the 32-bit limb ladders: base 2 by doublings, any other base by the binary ladder
    constexpr int L = ARITH == A_M2 ? 2 : (ARITH == A_M3 ? 3 : 4);
    constexpr bool FULL = ARITH == A_FULL;
    u32 x[L];
    if (b == 2) mont_pow2<L, FULL>(p, e, eb, x);
    else mont_powb<L, FULL>(p, e, eb, b, x);
    rlo = (u64)x[0] | ((u64)x[1] << 32);
    rhi = (L > 2 ? (u64)x[2] : 0) | (L > 3 ? ((u64)x[3] << 32) : 0);
  }
}

// Work item = one tile of up to 1024 primes of one segment, for one base.  Workgroups take
// items in turn and keep their sums until the (range, base) changes (uniform across the
// workgroup).  The bases' power tables are copied to LDS once per workgroup.
#ifndef M28_WAVES_EU
#define M28_WAVES_EU 8   // the radix-2^28 test kernel: 8 wavefronts per SIMD (<= 64 VGPRs; the
                         // ladder stays in registers, only its set-up spills a few values)
#endif
#ifndef F64_WAVES_EU
#define F64_WAVES_EU 6   // the FP64 test kernel: 6 wavefronts per SIMD (<= 80 VGPRs)
#endif
template <int ARITH>
__attribute__((amdgpu_waves_per_eu(ARITH == A_M28 ? M28_WAVES_EU : ARITH == A_F64 ? F64_WAVES_EU : 1)))
__global__ void __launch_bounds__(256) k_test(const SegDesc* __restrict__ segs, u32 nseg,
                                              const u32* __restrict__ off, const u32* __restrict__ cnt, u32 pcap,
                                              u32 tps, TestOut o) {
  __shared__ u32 pw_all[WF_MAXB * BPOW];
  for (u32 k = threadIdx.x; k < o.nb * BPOW; k += blockDim.x) pw_all[k] = c_bpow[k];
  __syncthreads();
  LaneSums s = {0, 0, 0, 0, 0};
  u32 cur = 0xFFFFFFFFu, curj = 0;
  const u64 nitems = (u64)nseg * tps * o.nb;
  for (u64 t = blockIdx.x; t < nitems; t += gridDim.x) {
    const u32 j = (u32)(t % o.nb);
    const u64 tile = t / o.nb;
    const u32 sg = (u32)(tile / tps);
    const u32 beg = (u32)(tile - (u64)sg * tps) * TEST_TILE;
    const u32 n = cnt[sg];
    if (beg >= n) continue;                       // uniform across the workgroup
    const u32 end = min(n, beg + TEST_TILE);
    const SegDesc d = segs[sg];
    if (d.rid != cur || j != curj) {
      if (cur != 0xFFFFFFFFu) flush_sums(o, cur, curj, s);
      cur = d.rid;
      curj = j;
    }
    const u32 b = c_base[j];
    const int w = (int)c_bwin[j];
    const u32* pw = pw_all + j * BPOW;
    const i64 nT = o.nearv ? o.nearv[d.rid] : o.near_T;
    const u32* po = off + (u64)sg * pcap;
    // exponent length of the segment's largest number: no prime of the tile has a longer
    // one, so the ladders of a wavefront agree without a cross-lane reduction
    const u64 etop = d.base + d.hi_off - 2;
    const int eb = etop ? 64 - __builtin_clzll(etop) : 1;
    if (ARITH == A_F64) {
      // two primes per lane: i and i + blockDim.x.  A segment at or above 2^32 uses the
      // shorter window step; at or above 2^40, 1/p comes from the segment's 1/base
      // (f64_recip_near).  A lane past the tile's end re-tests the tile's last prime, and its
      // result is not accounted.
      const bool big = d.base >= (1ull << 32);
      const bool near = d.base >= (1ull << 40);
      const double based = (double)d.base;     // p = base + offset < F64_PMAX: exact as a double
      const double ib = near ? 1.0 / based : 0.0;
      for (u32 base = beg; base < end; base += 2 * blockDim.x) {
        const u32 i0 = base + threadIdx.x, i1 = i0 + blockDim.x;
        const bool act0 = i0 < end, act1 = i1 < end;
        const u32 f0 = po[min(i0, end - 1)], f1 = po[min(i1, end - 1)];
        const u64 p0 = d.base + f0, p1 = d.base + f1;
        const double fd0 = (double)f0, fd1 = (double)f1;
        const double pd0 = based + fd0, pd1 = based + fd1;
        double ip0, ip1;
        if (near) {
          ip0 = f64_recip_near(ib, fd0, pd0);
          ip1 = f64_recip_near(ib, fd1, pd1);
        } else {
          ip0 = 1.0 / pd0;
          ip1 = 1.0 / pd1;
        }
        double a0, b0, a1, b1;
        if (b == 2 && big) {
          fp64_pow2x2_r<5, true>(pd0, ip0, p0 - 1, pd1, ip1, p1 - 1, eb, a0, b0, a1, b1);
        } else if (b == 2) {
          fp64_pow2x2_r<5>(pd0, ip0, p0 - 1, pd1, ip1, p1 - 1, eb, a0, b0, a1, b1);
        } else if (big) {
#define WF_POWC(W_, BIG_) fp64_powc_x2_r<W_, BIG_>(pd0, ip0, p0 - 1, pd1, ip1, p1 - 1, eb, pw, pw, a0, b0, a1, b1)
          switch (w) {
            case 5: WF_POWC(5, true); break;
            case 4: WF_POWC(4, true); break;
            case 3: WF_POWC(3, true); break;
            case 2: WF_POWC(2, true); break;
            default: WF_POWC(1, true); break;
          }
        } else {
          switch (w) {
            case 5: WF_POWC(5, false); break;
            case 4: WF_POWC(4, false); break;
            case 3: WF_POWC(3, false); break;
            case 2: WF_POWC(2, false); break;
            default: WF_POWC(1, false); break;
          }
#undef WF_POWC
        }
        if (act0 && !divides_base(p0, b)) account_padic(o, d.rid, j, p0, pd0, a0, b0, s, nT);
        if (act1 && !divides_base(p1, b)) account_padic(o, d.rid, j, p1, pd1, a1, b1, s, nT);
      }
    } else {
      for (u32 base = beg; base < end; base += blockDim.x) {
        const u32 i = base + threadIdx.x;
        const bool active = i < end;
        const u64 p = d.base + po[min(i, end - 1)];   // an idle lane re-tests the tile's last prime
        u64 rhi, rlo;
        residue<ARITH>(p, eb, b, w, pw, rhi, rlo);
        if (active && !divides_base(p, b)) account(o, d.rid, j, p, rhi, rlo, s, nT);
      }
    }
  }
  if (cur != 0xFFFFFFFFu) flush_sums(o, cur, curj, s);
}

// Composite-base mode (p < F64_PMAX): the quotients of the ladder primes l_s first, two ladders
// per lane on the lane's prime (fp64_powc_x2 with two power tables, window W for all), kept
// in LDS; then for each base j, q_j = sum of e_js q_s (mod p), since q_p(ab) = q_p(a) +
// q_p(b) (mod p) for p not dividing ab.  A pair whose ladder residue is not 1 mod p (not
// for a prime p) is computed directly as k_test does.  Sums per base are reduced per
// wavefront into LDS and flushed per range.
template <int W>
__attribute__((amdgpu_waves_per_eu(F64_WAVES_EU)))
__global__ void __launch_bounds__(256) k_test_derived(const SegDesc* __restrict__ segs, u32 nseg,
                                                      const u32* __restrict__ off, const u32* __restrict__ cnt,
                                                      u32 pcap, u32 tps, TestOut o, u32 nl) {
  __shared__ u32 lpw[WF_MAXL * BPOW];
  __shared__ u64 qs[WF_MAXL][256];                 // the lane's prime: q of ladder s, ~0 if r != 1
  __shared__ unsigned long long acks[WF_MAXB], asumq[WF_MAXB], atot[3];   // tests, near, flt
  for (u32 k = threadIdx.x; k < nl * BPOW; k += blockDim.x) lpw[k] = c_lpow[k];
  for (u32 k = threadIdx.x; k < o.nb; k += blockDim.x) acks[k] = asumq[k] = 0;
  if (threadIdx.x < 3) atot[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  u32 cur = 0xFFFFFFFFu;
  auto flush = [&]() {   // every thread, after a barrier
    for (u32 k = threadIdx.x; k < o.nb; k += blockDim.x) {
      if (acks[k] | asumq[k]) {
        atomicAdd(&o.cks[(u64)cur * o.nb + k], acks[k]);
        atomicAdd(&o.sumq[(u64)cur * o.nb + k], asumq[k]);
      }
      acks[k] = asumq[k] = 0;
    }
    if (threadIdx.x == 0) {
      DevAcc* a = &o.acc[cur];
      if (atot[0]) atomicAdd(&a->tests, atot[0]);
      if (atot[1]) atomicAdd(&a->near, atot[1]);
      if (atot[2]) atomicAdd(&a->flt, atot[2]);
      atot[0] = atot[1] = atot[2] = 0;
    }
  };
  const u64 nitems = (u64)nseg * tps;
  for (u64 t = blockIdx.x; t < nitems; t += gridDim.x) {
    const u32 sg = (u32)(t / tps);
    const u32 beg = (u32)(t - (u64)sg * tps) * TEST_TILE;
    const u32 n = cnt[sg];
    if (beg >= n) continue;                       // uniform across the workgroup
    const u32 end = min(n, beg + TEST_TILE);
    const SegDesc d = segs[sg];
    if (d.rid != cur) {
      __syncthreads();
      if (cur != 0xFFFFFFFFu) flush();
      __syncthreads();
      cur = d.rid;
    }
    const u32* po = off + (u64)sg * pcap;
    const u64 etop = d.base + d.hi_off - 2;
    const int eb = etop ? 64 - __builtin_clzll(etop) : 1;
    const bool big = d.base >= (1ull << 32);
    const u64 idle = big ? 4294967311ull : 3ull;
    const i64 nT = o.nearv ? o.nearv[d.rid] : o.near_T;
    const double based = (double)d.base;          // p = base + offset < F64_PMAX: exact as a double
    for (u32 base = beg; base < end; base += blockDim.x) {
      const u32 i = base + threadIdx.x;
      const bool act = i < end;
      const u32 f = act ? po[i] : 0;
      const u64 p = act ? d.base + f : idle;
      const double pd = act ? based + (double)f : (double)idle;
      for (u32 sl = 0; sl < nl; sl += 2) {
        const u32 s1 = sl + 1 < nl ? sl + 1 : sl;
        double a0, b0, a1, b1;
        if (big) fp64_powc_x2<W, true>(pd, p - 1, pd, p - 1, eb, lpw + sl * BPOW, lpw + s1 * BPOW, a0, b0, a1, b1);
        else fp64_powc_x2<W>(pd, p - 1, pd, p - 1, eb, lpw + sl * BPOW, lpw + s1 * BPOW, a0, b0, a1, b1);
        u64 q;
        qs[sl][threadIdx.x] = f64_quotient(a0, b0, pd, q) ? q : ~0ull;
        if (s1 != sl) qs[s1][threadIdx.x] = f64_quotient(a1, b1, pd, q) ? q : ~0ull;
      }
      for (u32 j = 0; j < o.nb; j++) {
        const u32 b = c_base[j];
        LaneSums ls = {0, 0, 0, 0, 0};
        if (act && !divides_base(p, b)) {
          u64 q = 0;
          bool ok = p != o.fault_p;         // an injected fault takes the direct ladder
          for (u32 sl = 0; sl < nl && ok; sl++) {
            const u32 ex = c_lexp[j * WF_MAXL + sl];
            if (!ex) continue;
            const u64 v = qs[sl][threadIdx.x];
            ok = v != ~0ull;
            for (u32 k = 0; ok && k < ex; k++) {
              q += v;
              if (q >= p) q -= p;
            }
          }
          if (ok) {
            ls.tests = 1;
            ls.cks = d_mix64(q ^ (p * GOLD));
            ls.sumq = q;
            if (q == 0) {
              push_event(o, WF_EV_FOUND, cur, j, p, 0, 0, 1);
            } else if (nT > 0 && near_hit(p, q, nT)) {
              ls.near = 1;
              const u64 lo = q * p, hi = __umul64hi(q, p);
              const u64 rlo = lo + 1;
              push_event(o, WF_EV_NEAR, cur, j, p, q, hi + (rlo < lo), rlo);
            }
          } else {
            u64 rhi, rlo;
            residue<A_F64>(p, eb, b, (int)c_bwin[j], c_bpow + j * BPOW, rhi, rlo);
            account(o, cur, j, p, rhi, rlo, ls, nT);
          }
        }
        const u64 c = wave_sum(ls.cks), sq = wave_sum(ls.sumq);
        const u64 tn = __popcll(__ballot(ls.tests != 0)), nn = __popcll(__ballot(ls.near != 0)),
                  fn = __popcll(__ballot(ls.flt != 0));
        if (lane == 0) {
          if (c | sq) {
            atomicAdd(&acks[j], (unsigned long long)c);
            atomicAdd(&asumq[j], (unsigned long long)sq);
          }
          if (tn) atomicAdd(&atot[0], (unsigned long long)tn);
          if (nn) atomicAdd(&atot[1], (unsigned long long)nn);
          if (fn) atomicAdd(&atot[2], (unsigned long long)fn);
        }
      }
    }
  }
  __syncthreads();
  if (cur != 0xFFFFFFFFu) flush();
}

// Quotients of arbitrary odd p (check / selftest) for the base in slot `slot` of the base
// tables: per-lane choice of the path, the scan's classes (int32 = 1 selects the 32-bit
// Montgomery ladders from F64_PMAX up).
__global__ void k_quot(const u64* __restrict__ p_in, u64* __restrict__ q_out, u64* __restrict__ rhi_out,
                       u64* __restrict__ rlo_out, u64 n, int int32, u32 slot) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u64 p = p_in[i];
  const int eb = 64 - __builtin_clzll(p - 1);
  const u32 b = c_base[slot];
  const int w = (int)c_bwin[slot];
  const u32* pw = c_bpow + slot * BPOW;
  u64 rhi, rlo;
  if (p < F64_PMAX) residue<A_F64>(p, eb, b, w, pw, rhi, rlo);
  else if (!int32) residue<A_M28>(p, eb, b, w, pw, rhi, rlo);
  else if (p < (1ull << 63)) residue<A_M4>(p, eb, b, w, pw, rhi, rlo);
  else residue<A_FULL>(p, eb, b, w, pw, rhi, rlo);
  q_out[i] = (rlo - 1) * inv64(p);
  rhi_out[i] = rhi;
  rlo_out[i] = rlo;
}

// Base-2 strong pseudoprimes, per tile of PSP_TILE odd numbers: the order sieve rules out
// odd multiples of 101 <= p < PSP_ORDER_MAX off their order class; the odd composites left
// (wheel composites from the sieve bitmap, and the odd multiples of 3 and 5) are listed in
// LDS; those of the list that pass the small-factor condition psp_possible are queued; the
// workgroup then runs the exact SPRP-2 test on the queue with every lane busy, and the n
// that pass are appended to out[] and counted per range.
#ifndef PSP_TILE
#define PSP_TILE 2048
#endif
#ifndef PSP_WAVES_EU
#define PSP_WAVES_EU 6
#endif
static_assert((u64)PSP_M4 + 2u * PSP_TILE < (1ull << 32) && 2u * PSP_TILE < PSP_M2,
              "a tile's offsets 2 li stay below PSP_M1..PSP_M4 (the smallest is PSP_M2) and keep r + 2 li in 32 bits");
#ifndef PSP_ORDER_MAX
#define PSP_ORDER_MAX 16384  // the order sieve takes the primes 101 <= p < PSP_ORDER_MAX (< 2^15)
#endif


__attribute__((amdgpu_waves_per_eu(PSP_WAVES_EU)))
__global__ void __launch_bounds__(256) k_psp(const SegDesc* __restrict__ segs, u32 nseg,
                                             const u32* __restrict__ bitmap, u32 bmw, u32 tps,
                                             const PspOrder* __restrict__ po, u32 npo,
                                             u64* __restrict__ out, u64 cap, unsigned long long* __restrict__ count,
                                             DevAcc* __restrict__ acc, unsigned long long* __restrict__ ntests) {
  __shared__ u64 queue[PSP_TILE];
  __shared__ unsigned short surv[PSP_TILE];   // odd composites the order sieve leaves (tile index)
  __shared__ u32 ruled[PSP_TILE / 32];   // bit i: odd number i of the tile ruled out by the order sieve
  __shared__ u32 qn, sn;
  __shared__ signed char w30[32];
  if (threadIdx.x < 30) w30[threadIdx.x] = c_W30[threadIdx.x];   // ordered by the first barrier below
  const int lane = threadIdx.x & 63;
  const u64 ntiles = (u64)nseg * tps;
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const u32 sg = (u32)(t / tps);
    const u32 k = (u32)(t - (u64)sg * tps);
    const SegDesc d = segs[sg];
    const u32 nodd = 15u * d.nbytes;                 // odd offsets 1, 3, ..., 30 nbytes - 1
    if (k * PSP_TILE >= nodd) continue;              // uniform across the workgroup
    if (threadIdx.x == 0) qn = sn = 0;
    for (u32 w = threadIdx.x; w < PSP_TILE / 32; w += blockDim.x) ruled[w] = 0;
    __syncthreads();
    // order sieve over the tile's odd numbers (psp_order_walk per prime)
    {
      const u64 A = d.base + 2ull * k * PSP_TILE + 1;
      for (u32 f = threadIdx.x; f < npo; f += blockDim.x)
        psp_order_walk(A, po[f], PSP_TILE, [&](u32 i) { atomicOr(&ruled[i >> 5], 1u << (i & 31u)); });
    }
    __syncthreads();
    // one LDS atomic per wavefront reserves the list slots of its lanes' entries
    auto push = [&](bool want, u32& cnt, auto store) {
      const u64 mask = __ballot(want);
      if (mask) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        u32 base = 0;
        if (lane == leader) base = atomicAdd(&cnt, (u32)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (want) store(base + __popcll(mask & ((1ull << lane) - 1)));
      }
    };
    // the odd composites left by the order sieve, compacted, so that the filter below runs
    // with every lane busy (about half the tile is ruled out)
    for (u32 j0 = 0; j0 < PSP_TILE; j0 += blockDim.x) {
      const u32 li = j0 + threadIdx.x;
      const u32 i = k * PSP_TILE + li;
      const u32 off = 2u * i + 1u;
      bool keep = false;
      if (i < nodd && off >= d.lo_off && off < d.hi_off && !((ruled[li >> 5] >> (li & 31u)) & 1u)) {
        const int cls = w30[off % 30u];
        keep = cls < 0 ? d.base + off > 5
                       : ((bitmap[(u64)sg * bmw + off / 120u] >> (((off / 30u) & 3u) * 8u + (u32)cls)) & 1u);
      }
      push(keep, sn, [&](u32 slot) { surv[slot] = (unsigned short)li; });
    }
    __syncthreads();
    const u32 ns = sn;
    // the tile's numbers are A + 2 li: their residues for psp_possible are A's (once per
    // tile) plus 2 li < 2^12, reduced by one subtraction
    const u64 A = d.base + 2ull * k * PSP_TILE + 1;
    const u32 r1 = (u32)(A % PSP_M1), r2 = (u32)(A % PSP_M2), r3 = (u32)(A % PSP_M3), r4 = (u32)(A % PSP_M4);
    const u32 r5 = (u32)(A % 97u);
    for (u32 j0 = 0; j0 < ns; j0 += blockDim.x) {    // uniform bound: every lane reaches the ballot
      const u32 j = j0 + threadIdx.x;
      u64 n = 0;
      bool want = false;
      if (j < ns) {
        const u32 o = 2u * surv[j];
        n = A + o;
        const u32 g1 = r1 + o, g2 = r2 + o, g3 = r3 + o, g4 = r4 + o;
        want = psp_possible_res((u32)n & 15u, g1 >= PSP_M1 ? g1 - PSP_M1 : g1, g2 >= PSP_M2 ? g2 - PSP_M2 : g2,
                                g3 >= PSP_M3 ? g3 - PSP_M3 : g3, g4 >= PSP_M4 ? g4 - PSP_M4 : g4, (r5 + o) % 97u);
      }
      push(want, qn, [&](u32 slot) { queue[slot] = n; });
    }
    __syncthreads();
    const u32 m = qn;
    if (threadIdx.x == 0 && m) atomicAdd(ntests, (unsigned long long)m);
    auto record = [&](u64 n) {
      const unsigned long long kk = atomicAdd(count, 1ull);
      if (kk < cap) out[kk] = n;
      atomicAdd(&acc[d.rid].npsp, 1ull);
    };
    for (u32 j = threadIdx.x; j < m; j += blockDim.x) {
      const u64 n = queue[j];
      if (sprp2_64(n)) record(n);
    }
    __syncthreads();   // the queue is refilled by the next tile
  }
}

// ---------------------------------------------------------------------------------
// prime statistics for the other sieve consumers (twin pairs, runs of increasing gaps)
// ---------------------------------------------------------------------------------
//
// The segment lists of a pass are ascending and the segments consecutive, so the primes
// after p are read straight from them.  A lane takes one prime p of [a, b) and walks the
// following primes while the gaps increase: L(p) = the number of increasing gaps
// (A158939/a158939.c:19-24); the first gap also gives the twin flag (A007508) and the gap
// maximum.  The pass sieves STATS_OV numbers past b so that every walk ends inside it.
#define STATS_OV (1ull << 20)
#define STATS_MAXRUN 64

struct StatsOut {
  unsigned long long primes, twins, gapkey;   // gapkey: gap << 48 | (2^48 - 1 - (p - a)), maximised
  unsigned long long runs[STATS_MAXRUN];
  unsigned long long first[STATS_MAXRUN];
};

__global__ void __launch_bounds__(256) k_stats(const SegDesc* __restrict__ segs, u32 nseg, const u32* __restrict__ off,
                                               const u32* __restrict__ cnt, u32 pcap, u32 tps, u64 a, u64 b,
                                               StatsOut* __restrict__ out, u32* __restrict__ flags) {
  __shared__ unsigned long long hist[STATS_MAXRUN], first[STATS_MAXRUN];
  if (threadIdx.x < STATS_MAXRUN) {
    hist[threadIdx.x] = 0;
    first[threadIdx.x] = ~0ull;
  }
  __syncthreads();
  const u64 M48 = (1ull << 48) - 1;
  u64 np = 0, ntw = 0, gk = 0;
  const u64 nitems = (u64)nseg * tps;
  for (u64 t = blockIdx.x; t < nitems; t += gridDim.x) {
    const u32 s = (u32)(t / tps);
    const u32 i = (u32)(t - (u64)s * tps) * 256u + threadIdx.x;
    if (i >= cnt[s]) continue;
    const u64 p = segs[s].base + off[(u64)s * pcap + i];
    if (p < a || p >= b) continue;
    np++;
    u32 cs = s, ci = i;
    u64 prev = p, gprev = 0;
    int L = 0;
    bool ended = false;
    for (int step = 0; step < STATS_MAXRUN; step++) {
      ci++;
      while (cs < nseg && ci >= cnt[cs]) {          // the next prime, in a later segment if need be
        cs++;
        ci = 0;
      }
      if (cs >= nseg) break;                          // past the sieved overlap: reported below
      const u64 q = segs[cs].base + off[(u64)cs * pcap + ci];
      const u64 g = q - prev;
      if (step == 0) {
        ntw += g == 2;
        gk = max(gk, (g << 48) | (M48 - (p - a)));
      }
      if (g <= gprev) {
        ended = true;
        break;
      }
      L++;
      gprev = g;
      prev = q;
    }
    if (!ended) {
      atomicOr(flags, 4u);
      continue;
    }
    atomicAdd(&hist[L], 1ull);
    atomicMin(&first[L], (unsigned long long)p);
  }
  np = wave_sum(np);
  ntw = wave_sum(ntw);
  for (int o = 32; o > 0; o >>= 1) gk = max(gk, (u64)__shfl_xor((unsigned long long)gk, o, 64));
  if ((threadIdx.x & 63) == 0) {
    if (np) atomicAdd(&out->primes, (unsigned long long)np);
    if (ntw) atomicAdd(&out->twins, (unsigned long long)ntw);
    if (gk) atomicMax(&out->gapkey, (unsigned long long)gk);
  }
  __syncthreads();
  if (threadIdx.x < STATS_MAXRUN && hist[threadIdx.x]) {
    atomicAdd(&out->runs[threadIdx.x], hist[threadIdx.x]);
    atomicMin(&out->first[threadIdx.x], first[threadIdx.x]);
  }
}

// BPSW verdicts of arbitrary n (bits: 1 base-2 strong probable prime, 2 strong Lucas
// probable prime, 4 prime).
__global__ void k_bpsw(const u64* __restrict__ n_in, unsigned char* __restrict__ out, u64 n) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (unsigned char)bpsw_flags(n_in[i]);
}

// The strong Lucas test of every base-2 strong pseudoprime listed by k_psp (one range per
// pass with the list): the ones that pass are BPSW pseudoprimes, of which none exist below
// 2^64.  A kernel of its own keeps the Lucas code's registers out of k_psp.
__global__ void k_psp_lucas(const u64* __restrict__ list, const unsigned long long* __restrict__ count, u64 cap,
                            DevAcc* __restrict__ acc) {
  const u64 n = min((u64)*count, cap);
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    if (slprp_64(list[i])) atomicAdd(&acc[0].nbpsw, 1ull);
}

// Zero n words (a kernel, so that the batched path can record it into a HIP graph).
__global__ void k_zero(u32* __restrict__ p, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------------

static const u32 R30H[8] = {1, 7, 11, 13, 17, 19, 23, 29};

// Staging area of wf_quotients_static, the entry point that allocates nothing: p, q, r_hi
// and r_lo of up to WF_STATIC_QUOT_MAX values (8 MiB).  Each GPU the library runs on holds
// its own copy until the process ends.  In host code the name is only a symbol: the device
// address of a GPU's copy comes from hipGetSymbolAddress with that GPU current, taken once
// per device by wf_open (scratch_addr) and used wherever host code hands the array out.
static __device__ __align__(256) u64 scratch[4 * (size_t)WF_STATIC_QUOT_MAX];

#define WF_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static u64* g_scratch_ptr[WF_MAX_DEVICES];

// The device address of `device`'s scratch array (the device must be current).
static int scratch_addr(int device, u64** out) {
  if (device < 0 || device >= WF_MAX_DEVICES) return WF_E_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return WF_E_DEVICE;
    g_scratch_ptr[device] = (u64*)ptr;
  }
  *out = g_scratch_ptr[device];
  return WF_OK;
}

struct BufSet {   // bucket counters / records, prime offsets / counts of one pass
  u32* bcnt = nullptr;
  u64 bcnt_cap = 0;
  u32* bdata = nullptr;
  u64 bdata_cap = 0;
  u32* off = nullptr;
  u64 off_cap = 0;
  u32* cnt = nullptr;
  u64 cnt_cap = 0;
  u32* bitmap = nullptr;   // composite bitmap (strong-pseudoprime pass only)
  u64 bitmap_cap = 0;
  u32* tcnt = nullptr;     // primes per prove tile
  u64 tcnt_cap = 0;
};

struct wf_engine {
  int dev = 0;
  hipStream_t st = nullptr;
  hipDeviceProp_t prop;
  wf_config cfg;
  int sbshift = 15;
  int nb = 1;
  u32 bases[WF_MAXB] = {2};
  int nl = 0;                    // composite-base mode: ladder primes (0 = one ladder per base)
  int lw = 0;                    // and their common window
  unsigned long long* d_cks = nullptr;    // per range and base (checksums, sums of q)
  unsigned long long* d_sumq = nullptr;
  u64 cks_cap = 0, sumq_cap = 0;
  // sieving primes (odd primes >= 7, ascending) and prefix sums of 1/q
  std::vector<u32> h_sp;
  std::vector<double> h_inv_sum;
  // fill slices: (first prime index of a [Q, 2Q) group, segments per run) -> the slice
  // ends over the whole group; cleared whenever the prime table or the configuration changes
  std::map<std::pair<u32, u32>, std::vector<u32>> slices;
  u64 sp_lim = 0;
  u32* d_sp = nullptr;
  u64* d_spinv = nullptr;
  double* d_sprcp = nullptr;     // 1/q rounded to nearest (placement below 2^53)
  u32 tierA = 0, q1idx = 0;
  double surv_frac = 1.0;        // share of numbers with no prime factor below bucket_min
  size_t fill_hint = 0;          // fill groups of the last planned pass (a reserve size)
  // pass descriptors (all passes of a batch) and two sets of pass buffers
  SegDesc* d_segs = nullptr;
  u64 seg_cap = 0;
  FillRun* d_runs = nullptr;     // fill runs of the pass(es)
  u64 run_cap = 0;
  u32* d_fills = nullptr;        // fill plans (FillGroup words and slice tables)
  u64 fill_cap = 0;
  BufSet bs[2];
  u32* d_flags = nullptr;
  u32* d_pre = nullptr;          // presieve pattern, 4 copies shifted by 0..3 bytes
  u32 pre_words = 0;
  hipStream_t st2 = nullptr;     // test kernels of the batched pipeline
  hipEvent_t gev[5] = {};        // graph capture: fork, sieved[2], tested[2]
  DevAcc* d_acc = nullptr;
  u64 acc_cap = 0;
  DevEvent* d_ev = nullptr;
  u64 dev_ev_cap = 0;
  StatsOut* d_stats = nullptr;           // wf_sieve_stats totals
  unsigned long long* d_ctr = nullptr;   // [1] events, [2] strong pseudoprimes, [3] proved primes,
                                         // [4] strong tests run by k_psp
  u64* d_psp = nullptr;                  // strong pseudoprimes found
  u64 psp_cap = 0;
  PspOrder* d_porder = nullptr;          // the order sieve's primes (built on first use)
  u32 nporder = 0;
  hipEvent_t ev[4] = {};
  wf_timing last;
  struct Server* srv = nullptr;
  u64* scratch = nullptr;          // this device's static scratch array (wf_quotients_static)
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
    case WF_E_UNSUPPORTED: return "not supported by the GPU engine";
    case WF_E_OVERFLOW: return "internal buffer bound exceeded";
    case WF_E_BUSY: return "engine busy serving (no free slot, or wf_serve_stop not called yet)";
    case WF_E_LIMIT: return "more values than the fixed scratch area of wf_quotients_static holds";
    default: return "unknown error";
  }
}

static u64 isqrt_u64(u64 n) {
  u64 r = (u64)std::sqrt((double)n);
  while (r > 0 && (unsigned __int128)r * r > n) r--;
  while ((unsigned __int128)(r + 1) * (r + 1) <= n) r++;
  return r;
}

// Odd primes 7 <= q <= lim on the host (segmented odd-only sieve, several threads),
// lim <= 2^32.  Built once per engine and kept on the device.
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
  const u64 SEGN = 1u << 21;  // odd numbers per segment
  const u64 nsegs = (lim + 2 * SEGN) / (2 * SEGN);
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nsegs < 4 * nt) nt = 1;
  std::vector<std::vector<u32>> part(nt);
  auto work = [&](unsigned t) {
    std::vector<uint8_t> comp(SEGN);
    for (u64 sgi = t * nsegs / nt; sgi < (t + 1) * nsegs / nt; sgi++) {
      const u64 lo = 1 + sgi * 2 * SEGN;         // odd numbers lo, lo + 2, ...
      if (lo > lim) break;
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
        if (!comp[k] && n >= 7 && n <= lim) part[t].push_back((u32)n);
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& v : part) total += v.size();
  out.reserve(total);
  for (auto& v : part) out.insert(out.end(), v.begin(), v.end());
}

// Primes >= cfg.bucket_min go through k_bucket_fill (index q1idx on); the numbers free of
// every smaller prime (the survivors of prove mode) are the share
// (8/30) prod_{7 <= q < bucket_min} (1 - 1/q) of all numbers.
static void set_bucket_min(wf_engine* e) {
  e->q1idx = (u32)(std::lower_bound(e->h_sp.begin(), e->h_sp.end(), e->cfg.bucket_min) - e->h_sp.begin());
  double f = 8.0 / 30.0;
  for (u32 i = 0; i < e->q1idx; i++) f *= 1.0 - 1.0 / e->h_sp[i];
  e->surv_frac = f;
  e->slices.clear();
}

static int ensure_sieving_primes(wf_engine* e, u64 lim) {
  // primes below 2^32 - 1 sieve every range below 2^64 (the table limit is capped there,
  // so asking for more must not rebuild it)
  if (lim > 0xFFFFFFFFull) lim = 0xFFFFFFFFull;
  if (lim <= e->sp_lim) return WF_OK;
  u64 want = std::max<u64>(lim, std::min<u64>(std::max<u64>(2 * e->sp_lim, 1u << 20), 0xFFFFFFFFull));
  if (want > 0xFFFFFFFFull) want = 0xFFFFFFFFull;
  host_primes(want, e->h_sp);
  std::vector<u64> inv(e->h_sp.size());
  for (size_t i = 0; i < e->h_sp.size(); i++) inv[i] = 0xFFFFFFFFFFFFFFFFull / e->h_sp[i];
  if (e->d_sp) (void)hipFree(e->d_sp);
  if (e->d_spinv) (void)hipFree(e->d_spinv);
  if (e->d_sprcp) (void)hipFree(e->d_sprcp);
  e->d_sprcp = nullptr;
  e->d_sp = nullptr;
  e->d_spinv = nullptr;
  const size_t n = std::max<size_t>(e->h_sp.size(), 1);
  if (hipMalloc(&e->d_sp, n * sizeof(u32)) != hipSuccess) return WF_E_NOMEM;
  if (hipMalloc(&e->d_spinv, n * sizeof(u64)) != hipSuccess) return WF_E_NOMEM;
  if (hipMalloc(&e->d_sprcp, n * sizeof(double)) != hipSuccess) return WF_E_NOMEM;
  HIPCHK(hipMemcpy(e->d_sp, e->h_sp.data(), e->h_sp.size() * sizeof(u32), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->d_spinv, inv.data(), inv.size() * sizeof(u64), hipMemcpyHostToDevice));
  {
    std::vector<double> rcp(e->h_sp.size());
    for (size_t i = 0; i < e->h_sp.size(); i++) rcp[i] = 1.0 / (double)e->h_sp[i];
    HIPCHK(hipMemcpy(e->d_sprcp, rcp.data(), rcp.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  e->sp_lim = want;
  // tier A: primes below 64 * 8 = 512 are striped over a wavefront
  e->tierA = (u32)(std::lower_bound(e->h_sp.begin(), e->h_sp.end(), 512u) - e->h_sp.begin());
  e->h_inv_sum.assign(e->h_sp.size() + 1, 0.0);
  for (size_t i = 0; i < e->h_sp.size(); i++) e->h_inv_sum[i + 1] = e->h_inv_sum[i] + 1.0 / e->h_sp[i];
  set_bucket_min(e);
  return WF_OK;
}

template <typename T>
static int grow(T*& ptr, u64& cap, u64 need) {
  if (need <= cap) return WF_OK;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  u64 n = std::max<u64>(need, cap + cap / 2);
  if (hipMalloc(&ptr, n * sizeof(T)) != hipSuccess) {
    cap = 0;
    return WF_E_NOMEM;
  }
  cap = n;
  return WF_OK;
}

static int put_base(u32 k, u32 b);

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
  if (!e->cfg.seg_bytes) e->cfg.seg_bytes = 65536;
  if (!e->cfg.sieve_block) e->cfg.sieve_block = 1024;
  if (!e->cfg.test_block) e->cfg.test_block = 256;
  if (!e->cfg.chunk) e->cfg.chunk = 10000000000ull;
  if (!e->cfg.bucket_min) e->cfg.bucket_min = 262144;
  if (!e->cfg.test_wgs) e->cfg.test_wgs = 16;
  if (e->cfg.test_wgs > 16 || e->cfg.seg_bytes < 4096 || (e->cfg.seg_bytes & (e->cfg.seg_bytes - 1)) ||
      e->cfg.seg_bytes > SEG_MAX || e->cfg.seg_bytes + 4096 > e->prop.sharedMemPerBlock || e->cfg.sieve_block < 64 ||
      e->cfg.sieve_block > 1024 || (e->cfg.sieve_block & 63) || e->cfg.test_block != 256 || e->cfg.arith > 1 ||
      e->cfg.bucket_min < 512 || e->cfg.chunk < 30ull * e->cfg.seg_bytes || e->cfg.prove > WF_PROVE_ON ||
      e->cfg.derive > WF_DERIVE_OFF || e->cfg.pipeline > WF_PIPELINE_OFF) {
    delete e;
    return WF_E_ARG;
  }
  e->sbshift = __builtin_ctz(e->cfg.seg_bytes);
  if (scratch_addr(device, &e->scratch) != WF_OK) {
    delete e;
    return WF_E_DEVICE;
  }
  // segments above 64 KiB need the kernel's dynamic LDS limit raised
  if (hipFuncSetAttribute((const void*)k_sieve, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)std::min<size_t>(e->prop.sharedMemPerBlock - 4096, SEG_MAX)) != hipSuccess) {
    (void)hipGetLastError();
    if (e->cfg.seg_bytes > 65536) {
      delete e;
      return WF_E_UNSUPPORTED;
    }
  }
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking) != hipSuccess) {
    wf_close(e);
    return WF_E_DEVICE;
  }
  for (int i = 0; i < 4; i++)
    if (hipEventCreate(&e->ev[i]) != hipSuccess) { wf_close(e); return WF_E_DEVICE; }
  for (int i = 0; i < 5; i++)
    if (hipEventCreateWithFlags(&e->gev[i], hipEventDisableTiming) != hipSuccess) { wf_close(e); return WF_E_DEVICE; }
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
  // wheel walk: for q = R30[k] (mod 30) and t = q m (mod 30), the first j >= 0 with m + j
  // prime to 30 and the index of (m + j) mod 30 in R30; the class of q R30[i] for each i
  unsigned char ws[8 * 30];
  u32 wcls[8];
  for (int k = 0; k < 8; k++) {
    const u32 q = R30H[k];
    u32 qinv = 1;
    while (qinv * q % 30 != 1) qinv++;
    for (u32 t = 0; t < 30; t++) {
      const u32 m = t * qinv % 30;
      u32 j = 0;
      while ((m + j) % 2 == 0 || (m + j) % 3 == 0 || (m + j) % 5 == 0) j++;
      ws[k * 30 + t] = (unsigned char)((j << 3) | qi[(m + j) % 30]);
    }
    wcls[k] = 0;
    for (int i = 0; i < 8; i++) wcls[k] |= (u32)qi[q * R30H[i] % 30] << (4 * i);
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_M30), m30, sizeof m30) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_QI), qi, sizeof qi) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_WS), ws, sizeof ws) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_WCLS), wcls, sizeof wcls) != hipSuccess || put_base(0, 2) != WF_OK) {
    wf_close(e);
    return WF_E_DEVICE;
  }
  // presieve pattern: byte J has bit c set iff 30 J + R30[c] is divisible by 7, 11, 13, 17 or 19;
  // copy k (k = 0..3) holds the words starting at byte 4 i + k, long enough for any phase
  // plus the largest segment
  {
    const u32 ext = PRE_PERIOD + SEG_MAX + 8;
    std::vector<unsigned char> pat(ext);
    for (u32 j = 0; j < ext; j++) {
      const u64 J = j % PRE_PERIOD;
      unsigned char b = 0;
      for (int c = 0; c < 8; c++) {
        const u64 n = 30 * J + R30H[c];
        if (n % 7 == 0 || n % 11 == 0 || n % 13 == 0 || n % 17 == 0 || n % 19 == 0) b |= (unsigned char)(1u << c);
      }
      pat[j] = b;
    }
    e->pre_words = (PRE_PERIOD + SEG_MAX) / 4 + 1;
    std::vector<u32> w4(4 * (size_t)e->pre_words);
    for (u32 k = 0; k < 4; k++)
      for (u32 i = 0; i < e->pre_words; i++) {
        const u32 o = 4 * i + k;
        w4[(size_t)k * e->pre_words + i] = (u32)pat[o] | ((u32)pat[o + 1] << 8) | ((u32)pat[o + 2] << 16) |
                                          ((u32)pat[o + 3] << 24);
      }
    if (hipMalloc(&e->d_pre, w4.size() * sizeof(u32)) != hipSuccess ||
        hipMemcpy(e->d_pre, w4.data(), w4.size() * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess) {
      wf_close(e);
      return WF_E_NOMEM;
    }
  }
  if (hipMalloc(&e->d_ctr, 8 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&e->d_flags, 8 * sizeof(u32)) != hipSuccess) {
    wf_close(e);
    return WF_E_NOMEM;
  }
  memset(&e->last, 0, sizeof e->last);
  *out = e;
  return WF_OK;
}

extern "C" void wf_close(wf_engine* e) {
  if (!e) return;
  if (e->srv) (void)wf_serve_stop(e);
  (void)hipSetDevice(e->dev);
  if (e->st) (void)hipStreamSynchronize(e->st);
  (void)hipFree(e->d_sp);
  (void)hipFree(e->d_spinv);
  (void)hipFree(e->d_sprcp);
  (void)hipFree(e->d_segs);
  (void)hipFree(e->d_runs);
  (void)hipFree(e->d_fills);
  for (int k = 0; k < 2; k++) {
    (void)hipFree(e->bs[k].bcnt);
    (void)hipFree(e->bs[k].bdata);
    (void)hipFree(e->bs[k].off);
    (void)hipFree(e->bs[k].cnt);
    (void)hipFree(e->bs[k].bitmap);
    (void)hipFree(e->bs[k].tcnt);
  }
  (void)hipFree(e->d_psp);
  (void)hipFree(e->d_porder);
  (void)hipFree(e->d_stats);
  (void)hipFree(e->d_pre);
  (void)hipFree(e->d_flags);
  (void)hipFree(e->d_acc);
  (void)hipFree(e->d_cks);
  (void)hipFree(e->d_sumq);
  (void)hipFree(e->d_ev);
  (void)hipFree(e->d_ctr);
  for (int i = 0; i < 4; i++)
    if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
  for (int i = 0; i < 5; i++)
    if (e->gev[i]) (void)hipEventDestroy(e->gev[i]);
  if (e->st) (void)hipStreamDestroy(e->st);
  if (e->st2) (void)hipStreamDestroy(e->st2);
  delete e;
}

extern "C" int wf_info(wf_engine* e, char* buf, int len) {
  if (!e || !buf || len <= 0) return WF_E_ARG;
  snprintf(buf, (size_t)len, "%s (%s), %d CUs, LDS segment %u B (%zu B per workgroup), %s", e->prop.name,
           e->prop.gcnArchName, e->prop.multiProcessorCount, e->cfg.seg_bytes, (size_t)e->prop.sharedMemPerBlock,
           e->cfg.arith == 1 ? "32-bit Montgomery" : "FP64 p-adic below 9*2^47, radix-2^28 Montgomery above");
  return WF_OK;
}

// Window and powers of base b: the largest w <= 5 with b^(2^w - 1) < 2^32 (the rule of
// wieferich.c's bases_setup, A001220/wieferich.c:99-107) and pw[d] = b^d for d < 2^w.
static int base_window(u32 b, u32* pw) {
  int w = 5;
  while (w > 1) {
    unsigned __int128 x = 1;
    bool fits = true;
    for (int k = 1; k < (1 << w) && fits; k++) {
      x *= b;
      fits = (x >> 32) == 0;
    }
    if (fits) break;
    w--;
  }
  u64 x = 1;
  for (int d = 0; d < BPOW; d++) {
    pw[d] = d < (1 << w) ? (u32)x : 0;
    x *= b;
  }
  return w;
}

// Upload base b's table into slot k of the constant tables.
static int put_base(u32 k, u32 b) {
  u32 pw[BPOW];
  const u32 w = (u32)base_window(b, pw);
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_base), &b, sizeof b, k * sizeof(u32)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_bwin), &w, sizeof w, k * sizeof(u32)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_bpow), pw, sizeof pw, (size_t)k * BPOW * sizeof(u32)) != hipSuccess)
    return WF_E_DEVICE;
  return WF_OK;
}

// Composite-base mode for a base list: the distinct primes dividing the bases become the
// ladders when they are fewer than the bases (at most WF_MAXL), with one window for all
// (the smallest of theirs) and each base's exponent of each.  nl = 0 otherwise.
static int plan_ladders(wf_engine* e) {
  e->nl = 0;
  if (e->cfg.derive == WF_DERIVE_OFF || e->cfg.arith != 0) return WF_OK;
  std::vector<u32> ls;
  std::vector<std::vector<std::pair<u32, u32>>> fac(e->nb);   // (prime, exponent) per base
  for (int j = 0; j < e->nb; j++) {
    u32 x = e->bases[j];
    for (u32 f = 2; (u64)f * f <= x; f++) {
      u32 k = 0;
      while (x % f == 0) { x /= f; k++; }
      if (k) fac[j].push_back({f, k});
    }
    if (x > 1) fac[j].push_back({x, 1});
    for (auto& fk : fac[j]) ls.push_back(fk.first);
  }
  std::sort(ls.begin(), ls.end());
  ls.erase(std::unique(ls.begin(), ls.end()), ls.end());
  if ((int)ls.size() >= e->nb || ls.size() > WF_MAXL) return WF_OK;
  int w = 5;
  u32 pw[BPOW];
  for (u32 l : ls) w = std::min(w, base_window(l, pw));
  std::vector<u32> lpow(WF_MAXL * BPOW, 0);
  for (size_t s = 0; s < ls.size(); s++) {
    base_window(ls[s], pw);                  // l^d for d below its own window, at least ours
    for (int d = 0; d < (1 << w); d++) lpow[s * BPOW + d] = pw[d];
  }
  std::vector<unsigned char> lexp(WF_MAXB * WF_MAXL, 0);
  for (int j = 0; j < e->nb; j++)
    for (auto& fk : fac[j])
      lexp[j * WF_MAXL + (std::lower_bound(ls.begin(), ls.end(), fk.first) - ls.begin())] = (unsigned char)fk.second;
  std::vector<u32> lb(WF_MAXL, 0);
  std::copy(ls.begin(), ls.end(), lb.begin());
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_lbase), lb.data(), WF_MAXL * sizeof(u32)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_lpow), lpow.data(), lpow.size() * sizeof(u32)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_lexp), lexp.data(), lexp.size()) != hipSuccess)
    return WF_E_DEVICE;
  e->nl = (int)ls.size();
  e->lw = w;
  return WF_OK;
}

extern "C" int wf_set_bases(wf_engine* e, const u32* bases, int nb) {
  if (!e || !bases || nb < 1 || nb > WF_MAXB) return WF_E_ARG;
  for (int i = 0; i < nb; i++)
    if (bases[i] < 2 || (i && bases[i] <= bases[i - 1])) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  HIPCHK(hipSetDevice(e->dev));
  for (int i = 0; i < nb; i++) {
    const int rc = put_base((u32)i, bases[i]);
    if (rc != WF_OK) return rc;
  }
  e->nb = nb;
  memcpy(e->bases, bases, (size_t)nb * sizeof(u32));
  return plan_ladders(e);
}

// A sub-range of one arithmetic class: [lo, hi) (hi == 0 means 2^64) with its kernel.
struct Piece { u64 lo, hi; u32 rid; int arith; };

// Split [lo, hi) at the class boundaries of the arithmetic paths.
static void split_classes(const wf_engine* e, u64 lo, u64 hi, u32 rid, std::vector<Piece>& out) {
  std::vector<std::pair<u64, int>> cuts;   // (upper bound of the class, class)
  if (e->cfg.arith == 1) {
    cuts = {{1ull << 31, A_M2}, {1ull << 47, A_M3}, {1ull << 63, A_M4}, {0, A_FULL}};
  } else {
    cuts = {{F64_PMAX, A_F64}, {0, A_M28}};
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

// Segment descriptors of [lo, hi) and the runs (<= RUN_MAX consecutive segments) they form.
static void build_segments(const wf_engine* e, u64 lo, u64 hi, u32 rid, bool prove, std::vector<SegDesc>& segs,
                           std::vector<RunDesc>& runs) {
  const u32 SB = e->cfg.seg_bytes;
  const u64 span = 30ull * SB;
  u64 base = lo - lo % 30;
  // isqrt of the segment's last number and the count of sieving primes up to it: both grow
  // with the segment, so each advances from the previous segment's value
  u64 r = 0;
  u32 nq = 0;
  bool first = true;
  const u32 nsp = (u32)e->h_sp.size();
  for (;;) {
    SegDesc d;
    d.base = base;
    d.rid = rid;
    const u64 room = 0 - base;  // 2^64 - base (0 if base == 0: unlimited)
    u64 len = span;
    if (base != 0 && room < len) len = room;
    if (hi != 0 && hi - base < len) len = hi - base;
    const u32 nbytes = (u32)((len + 29) / 30);
    d.nbytes = (nbytes + 3) & ~3u;
    d.lo_off = lo > base ? (u32)(lo - base) : 0;
    if (base == 0 && d.lo_off < 2) d.lo_off = 2;   // 1 is not prime
    d.hi_off = (u32)len;
    const u64 last = base + (len - 1);
    if (first) {
      first = false;
      r = isqrt_u64(last);
      nq = (u32)(std::upper_bound(e->h_sp.begin(), e->h_sp.end(), (u32)std::min<u64>(r, 0xFFFFFFFFull)) -
                 e->h_sp.begin());
    } else {
      while ((unsigned __int128)(r + 1) * (r + 1) <= last) r++;
      const u32 rq = (u32)std::min<u64>(r, 0xFFFFFFFFull);
      while (nq < nsp && e->h_sp[nq] <= rq) nq++;
    }
    d.nq = nq;
    d.prove = 0;
    if (prove && d.nq > e->q1idx) {   // sieve with the primes below bucket_min, prove the rest
      d.nq = e->q1idx;
      d.prove = 1;
    }
    if (runs.empty() || runs.back().seg0 + runs.back().nseg != segs.size() || runs.back().nseg == RUN_MAX ||
        segs.empty() || segs.back().rid != rid || segs.back().base + span != base) {
      RunDesc rd;
      rd.base = base;
      rd.seg0 = (u32)segs.size();
      rd.nseg = 0;
      rd.nq = 0;
      rd.pad = 0;
      runs.push_back(rd);
    }
    runs.back().nseg++;
    runs.back().nq = std::max(runs.back().nq, d.nq);
    segs.push_back(d);
    if (len < span) break;
    base += span;
    if (hi != 0 && base >= hi) break;
    if (base == 0) break;
  }
}

struct PassPlan {
  std::vector<SegDesc> segs;
  std::vector<RunDesc> runs;
  std::vector<FillRun> fruns;
  std::vector<u32> fills;          // the fill plan: groups, then slice tables (FillGroup)
  u32 nfgroups = 0, nfitems = 0;   // fill groups and k_bucket_fill work items
  u32 maxnq = 0, bcap = 0, pcap = 0, tps = 0, nprove = 0;
};

// Prove-mode cost model, in nanoseconds of the whole device, measured on MI355X: placing
// one sieving prime in one piece (k_bucket_fill: 3 to 5 ps from 1e17 to 2^64), BPSW per
// survivor (base-2 strong test, plus the strong Lucas test for the half that are prime:
// 0.24 ns), doubled because the bucket fill overlaps the tests of the batched path and
// BPSW, all VALU work, does not, and the prove kernels' own start (0.08 ms).
#define PROVE_T_PLACE 0.0035
#define PROVE_T_SURV 0.48
#define PROVE_T_FIXED 80000.0

// Whether piece [lo, hi) should run in prove mode: its sieving primes from bucket_min up
// would each be placed at least once, against BPSW for its survivors.
static bool prove_piece(const wf_engine* e, u64 lo, u64 hi) {
  if (e->cfg.prove == WF_PROVE_OFF) return false;
  const u64 top = hi ? hi - 1 : 0xFFFFFFFFFFFFFFFFull;
  const u32 nq = (u32)(std::upper_bound(e->h_sp.begin(), e->h_sp.end(),
                                        (u32)std::min<u64>(isqrt_u64(top), 0xFFFFFFFFull)) - e->h_sp.begin());
  if (nq <= e->q1idx) return false;   // nothing to leave out
  if (e->cfg.prove == WF_PROVE_ON) return true;
  return span_d(lo, hi) * e->surv_frac * PROVE_T_SURV + PROVE_T_FIXED < (double)(nq - e->q1idx) * PROVE_T_PLACE;
}

// Entries per segment list: the primes of a segment, or in prove mode its survivors (their
// expected count E = surv_frac numbers, with room for the spread of a sieved count).
static u32 list_cap(const wf_engine* e, bool any_prove) {
  const double y = 30.0 * e->cfg.seg_bytes;
  u64 cap = prime_bound(y);
  if (any_prove) {
    const double E = y * e->surv_frac;
    cap = std::max<u64>(cap, (u64)(1.25 * E + 8.0 * std::sqrt(E) + 512.0));
  }
  return (u32)((cap + 63) & ~63ull);
}

// Slice ends for the primes [i, gend) of one [Q, 2Q) group over runs of ns segments: each
// slice holds about FILL_TARGET expected records (prime q leaves 8 seg_bytes ns / q of them),
// and at most 65536 primes.  Cached per (group, ns): passes of one range repeat them.
static const std::vector<u32>& fill_slices(wf_engine* e, u32 i, u32 gend, u32 ns) {
  auto key = std::make_pair(i, ns);
  auto it = e->slices.find(key);
  if (it != e->slices.end()) return it->second;
  std::vector<u32> ends;
  const double per = 8.0 * e->cfg.seg_bytes * ns;
  u32 a = i;
  while (a < gend) {
    const double lim = e->h_inv_sum[a] + FILL_TARGET / per;
    u32 b = (u32)(std::upper_bound(e->h_inv_sum.begin() + a + 1, e->h_inv_sum.begin() + gend + 1, lim) -
                  e->h_inv_sum.begin()) - 1;
    if (b <= a) b = a + 1;
    b = std::min(b, std::min<u32>(gend, a + 65536));
    ends.push_back(b);
    a = b;
  }
  return e->slices.emplace(key, std::move(ends)).first->second;
}

static void plan_pass(wf_engine* e, const std::vector<Piece>& pieces, PassPlan& P, bool allow_prove = true) {
  P.segs.clear();
  P.runs.clear();
  P.fruns.clear();
  for (const Piece& p : pieces)
    build_segments(e, p.lo, p.hi, p.rid, allow_prove && prove_piece(e, p.lo, p.hi), P.segs, P.runs);
  P.maxnq = 0;
  for (const RunDesc& r : P.runs) P.maxnq = std::max(P.maxnq, r.nq);
  P.nprove = 0;
  for (const SegDesc& d : P.segs) P.nprove += d.prove;
  P.pcap = list_cap(e, P.nprove > 0);
  P.tps = (P.pcap + TEST_TILE - 1) / TEST_TILE;
  P.bcap = 0;
  P.fills.clear();
  P.nfgroups = 0;
  P.nfitems = 0;
  std::vector<FillGroup> groups;
  groups.reserve(e->fill_hint);
  std::vector<u32> tables;                                // slice tables of the pass, each once
  std::vector<std::pair<const std::vector<u32>*, u32>> toffs;   // table -> its offset in tables
  if (P.maxnq > e->q1idx) {
    // expected records per segment: sum over large primes and classes of seg_bytes / q; the
    // count is a sum of independent 0/1 (or floor/ceil) terms, so its deviation is ~sqrt(E)
    const double E = 8.0 * e->cfg.seg_bytes * (e->h_inv_sum[P.maxnq] - e->h_inv_sum[e->q1idx]);
    P.bcap = (u32)((E + 10.0 * std::sqrt(E) + 1024.0 + 63.0) / 64.0) * 64;
    // fill groups.  Primes in [Q, 2Q) use groups of at most G segments, G the power of two
    // (<= RUN_MAX) nearest 64 Q / (8 seg_bytes): about 64 records per prime and group, so a
    // prime's start in a group is paid for by its hits.  A group is a run of one range cut
    // to G segments, or several short runs (of short ranges) packed together.  Each
    // group's primes are cut into slices of ~FILL_TARGET expected records.
    const double SBd = e->cfg.seg_bytes;
    u32 i = e->q1idx;
    while (i < P.maxnq) {
      const u32 Q = e->h_sp[i];
      const u32 qend = Q >= 0x80000000u ? 0xFFFFFFFFu : 2 * Q;
      const u32 gend = (u32)(std::lower_bound(e->h_sp.begin() + i, e->h_sp.end(), qend) - e->h_sp.begin());
      u32 G = 1;
      while (G < RUN_MAX && 2.0 * G * 8.0 * SBd <= FILL_PER_PRIME * (double)Q) G *= 2;
      u32 g_fr0 = (u32)P.fruns.size(), g_nseg = 0, g_nq = 0;
      u64 g_top = 0;   // end of the group's highest run
      auto flush = [&]() {
        if (!g_nseg) return;
        const u32 bend = std::min(gend, g_nq);   // > i: every run of the group uses prime i
        const std::vector<u32>& ends = fill_slices(e, i, gend, g_nseg);
        u32 toff = ~0u;
        for (const auto& t : toffs)
          if (t.first == &ends) toff = t.second;
        if (toff == ~0u) {
          toff = (u32)tables.size();
          tables.insert(tables.end(), ends.begin(), ends.end());
          toffs.emplace_back(&ends, toff);
        }
        // the slices [i, ends[0]), [ends[0], ends[1]), ... that start below bend
        const u32 nsl = 1 + (u32)(std::lower_bound(ends.begin(), ends.end() - 1, bend) - ends.begin());
        FillGroup g;
        g.fr0 = g_fr0;
        g.nfr = (u32)P.fruns.size() - g_fr0;
        g.nseg = g_nseg;
        g.fp = g_top <= (1ull << 53) ? 1u : 0u;
        g.istart = i;
        g.bend = bend;
        g.slices = toff;       // relative to the tables until they are placed after the groups
        g.item0 = P.nfitems;
        groups.push_back(g);
        P.nfitems += nsl;
        g_fr0 = (u32)P.fruns.size();
        g_nseg = 0;
        g_nq = 0;
        g_top = 0;
      };
      for (const RunDesc& ch : P.runs) {
        if (ch.nq <= i) continue;
        for (u32 s0 = 0; s0 < ch.nseg; s0 += G) {
          const u32 ns = std::min(G, ch.nseg - s0);
          if (g_nseg + ns > G) flush();
          FillRun fr;
          fr.base = ch.base + (u64)s0 * 30ull * e->cfg.seg_bytes;
          fr.seg0 = ch.seg0 + s0;
          fr.nseg = ns;
          fr.lseg0 = g_nseg;
          fr.pad = 0;
          P.fruns.push_back(fr);
          g_nseg += ns;
          g_nq = std::max(g_nq, ch.nq);
          // (a base at or above 2^53 counts as the top: base + span could wrap near 2^64)
          g_top = std::max<u64>(g_top, fr.base >= (1ull << 53) ? ~0ull : fr.base + (u64)ns * 30ull * e->cfg.seg_bytes);
        }
      }
      flush();
      i = gend;
    }
  }
  P.nfgroups = (u32)groups.size();
  const u32 nidx = (P.nfitems + 63) / 64, tbase = 8 * P.nfgroups + nidx;
  for (FillGroup& g : groups) g.slices += tbase;
  P.fills.resize((size_t)tbase + tables.size());
  if (!groups.empty()) memcpy(P.fills.data(), groups.data(), groups.size() * sizeof(FillGroup));
  for (u32 k = 0, gi = 0; k < nidx; k++) {   // the group of item 64 k
    while (gi + 1 < P.nfgroups && groups[gi + 1].item0 <= 64 * k) gi++;
    P.fills[8 * P.nfgroups + k] = gi;
  }
  if (!tables.empty()) memcpy(P.fills.data() + tbase, tables.data(), tables.size() * sizeof(u32));
  e->fill_hint = groups.size();
}

// Prove tiles per segment list.
static u32 prove_tiles(const PassPlan& P) { return (P.pcap + PROVE_TILE - 1) / PROVE_TILE; }

// Make a buffer set large enough for pass P.
static int size_set(BufSet& B, const PassPlan& P) {
  const u64 nseg = P.segs.size();
  int rc;
  if ((rc = grow(B.cnt, B.cnt_cap, nseg)) != WF_OK) return rc;
  if ((rc = grow(B.off, B.off_cap, nseg * P.pcap)) != WF_OK) return rc;
  if (P.bcap) {
    if ((rc = grow(B.bcnt, B.bcnt_cap, nseg)) != WF_OK) return rc;
    if ((rc = grow(B.bdata, B.bdata_cap, nseg * (u64)P.bcap)) != WF_OK) return rc;
  }
  if (P.nprove && (rc = grow(B.tcnt, B.tcnt_cap, nseg * (u64)prove_tiles(P))) != WF_OK) return rc;
  return WF_OK;
}

// Bucket fill + sieve of pass P (descriptors at dsegs / druns) into buffer set B on stream st.
static int enqueue_sieve(wf_engine* e, const PassPlan& P, const SegDesc* dsegs, const FillRun* druns,
                         const u32* dfills, BufSet& B, hipStream_t st, DevAcc* acc, bool want_bitmap = false) {
  const u64 nseg = P.segs.size();
  const int cus = e->prop.multiProcessorCount;
  if (P.bcap) {
    hipLaunchKernelGGL(k_zero, dim3((u32)std::min<u64>((nseg + 255) / 256, 1024)), dim3(256), 0, st, B.bcnt, nseg);
    if (P.nfitems) {
      hipLaunchKernelGGL(k_bucket_fill, dim3(P.nfitems), dim3(FILL_BLOCK), 0, st, dfills, P.nfgroups, druns, e->d_sp,
                         e->d_spinv, e->d_sprcp, B.bcnt, B.bdata, P.bcap, e->sbshift, e->d_flags);
      HIPCHK(hipGetLastError());
    }
  }
  const u32 sgrid = (u32)std::min<u64>(nseg, (u64)cus * 4);
  hipLaunchKernelGGL(k_sieve, dim3(sgrid), dim3(e->cfg.sieve_block), e->cfg.seg_bytes, st, dsegs, (u32)nseg, e->d_sp,
                     e->d_spinv, e->d_sprcp, e->tierA, e->q1idx, B.bcnt, B.bdata, P.bcap, B.off, B.cnt, P.pcap, acc,
                     e->d_flags,
                     want_bitmap ? B.bitmap : nullptr, e->cfg.seg_bytes / 4, e->d_pre, e->pre_words);
  HIPCHK(hipGetLastError());
  if (P.nprove) {
    const u32 ptps = prove_tiles(P);
    hipLaunchKernelGGL(k_prove, dim3((u32)std::min<u64>((u64)P.nprove * ptps, (u64)cus * 16)), dim3(256), 0, st,
                       dsegs, (u32)nseg, ptps, e->d_sp, B.off, B.cnt, P.pcap, B.tcnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_prove_gather, dim3((u32)std::min<u64>(P.nprove, (u64)cus * 2)), dim3(PROVE_TILE), 0, st,
                       dsegs, (u32)nseg, ptps, B.off, B.cnt, P.pcap, B.tcnt, acc, e->d_ctr + 3);
    HIPCHK(hipGetLastError());
  }
  return WF_OK;
}

// Fermat-quotient tests of pass P's primes (buffer set B) on stream st.
// wgs = test workgroups per CU: a full machine (16) when the tests run alone, fewer in the
// pipelined path so that the next pass's sieve finds room beside them.
static int enqueue_test(wf_engine* e, const PassPlan& P, const SegDesc* dsegs, BufSet& B, hipStream_t st, int arith,
                        const TestOut& to, u32 wgs = 16) {
  const u64 nseg = P.segs.size();
  const u32 tgrid = (u32)std::min<u64>(nseg * P.tps, (u64)e->prop.multiProcessorCount * wgs);
  if (arith == A_F64 && e->nl > 0) {   // composite bases from the ladder primes' quotients
#define WF_TESTD(W) hipLaunchKernelGGL(k_test_derived<W>, dim3(tgrid), dim3(256), 0, st, dsegs, (u32)nseg, B.off, \
                                       B.cnt, P.pcap, P.tps, to, (u32)e->nl)
    switch (e->lw) {
      case 5: WF_TESTD(5); break;
      case 4: WF_TESTD(4); break;
      case 3: WF_TESTD(3); break;
      case 2: WF_TESTD(2); break;
      default: WF_TESTD(1); break;
    }
#undef WF_TESTD
    HIPCHK(hipGetLastError());
    return WF_OK;
  }
#define WF_TEST(A) hipLaunchKernelGGL(k_test<A>, dim3(tgrid), dim3(256), 0, st, dsegs, (u32)nseg, B.off, B.cnt, \
                                      P.pcap, P.tps, to)
  switch (arith) {
    case A_F64: WF_TEST(A_F64); break;
    case A_M2: WF_TEST(A_M2); break;
    case A_M3: WF_TEST(A_M3); break;
    case A_M4: WF_TEST(A_M4); break;
    case A_FULL: WF_TEST(A_FULL); break;
    default: WF_TEST(A_M28); break;
  }
#undef WF_TEST
  HIPCHK(hipGetLastError());
  return WF_OK;
}

// Upload the descriptors of pass P to the start of the descriptor arrays.
static int upload_pass(wf_engine* e, const PassPlan& P) {
  int rc;
  if ((rc = grow(e->d_segs, e->seg_cap, P.segs.size())) != WF_OK) return rc;
  if ((rc = grow(e->d_runs, e->run_cap, std::max<size_t>(P.fruns.size(), 1))) != WF_OK) return rc;
  HIPCHK(hipMemcpyAsync(e->d_segs, P.segs.data(), P.segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, e->st));
  if (!P.fruns.empty())
    HIPCHK(hipMemcpyAsync(e->d_runs, P.fruns.data(), P.fruns.size() * sizeof(FillRun), hipMemcpyHostToDevice, e->st));
  if (!P.fills.empty()) {
    if ((rc = grow(e->d_fills, e->fill_cap, P.fills.size())) != WF_OK) return rc;
    HIPCHK(hipMemcpyAsync(e->d_fills, P.fills.data(), P.fills.size() * sizeof(u32), hipMemcpyHostToDevice, e->st));
  }
  return WF_OK;
}

// Plain path: sieve and test one pass on the engine's stream, then wait for it.
static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The order sieve's table (k_psp): every prime 101 <= p < PSP_ORDER_MAX (psp_order_entry).
// Built once.
static int ensure_porder(wf_engine* e) {
  if (e->d_porder) return WF_OK;
  std::vector<PspOrder> t;
  for (u32 p = 101; p < PSP_ORDER_MAX; p += 2) {
    bool prime = true;
    for (u32 d = 3; d * d <= p; d += 2)
      if (p % d == 0) { prime = false; break; }
    if (prime) t.push_back(psp_order_entry(p));
  }
  if (hipMalloc(&e->d_porder, t.size() * sizeof(PspOrder)) != hipSuccess) {
    e->d_porder = nullptr;
    return WF_E_NOMEM;
  }
  HIPCHK(hipMemcpy(e->d_porder, t.data(), t.size() * sizeof(PspOrder), hipMemcpyHostToDevice));
  e->nporder = (u32)t.size();
  return WF_OK;
}

static int run_pieces(wf_engine* e, const std::vector<Piece>& pieces, int arith, const TestOut& to,
                      bool want_psp = false) {
  PassPlan P;
  const auto t0 = std::chrono::steady_clock::now();
  plan_pass(e, pieces, P, !want_psp);   // the pseudoprime pass needs the full composite bitmap
  e->last.plan_ms += ms_since(t0);
  if (P.segs.empty()) return WF_OK;
  int rc;
  if ((rc = size_set(e->bs[0], P)) != WF_OK) return rc;
  if (want_psp && (rc = grow(e->bs[0].bitmap, e->bs[0].bitmap_cap, P.segs.size() * (e->cfg.seg_bytes / 4))) != WF_OK)
    return rc;
  if (want_psp && (rc = ensure_porder(e)) != WF_OK) return rc;
  if ((rc = upload_pass(e, P)) != WF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev[0], e->st));
  if ((rc = enqueue_sieve(e, P, e->d_segs, e->d_runs, e->d_fills, e->bs[0], e->st, to.acc, want_psp)) != WF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev[1], e->st));
  if ((rc = enqueue_test(e, P, e->d_segs, e->bs[0], e->st, arith, to)) != WF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev[2], e->st));
  if (want_psp) {
    const u32 tps = (15u * e->cfg.seg_bytes + PSP_TILE - 1) / PSP_TILE;
    const u64 nt = P.segs.size() * (u64)tps;
    hipLaunchKernelGGL(k_psp, dim3((u32)std::min<u64>(nt, (u64)e->prop.multiProcessorCount * 16)), dim3(256), 0, e->st,
                       e->d_segs, (u32)P.segs.size(), e->bs[0].bitmap, e->cfg.seg_bytes / 4, tps, e->d_porder,
                       e->nporder, e->d_psp, e->psp_cap, e->d_ctr + 2, to.acc, e->d_ctr + 4);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e->ev[3], e->st));
  }
  u32 flags = 0;
  HIPCHK(hipMemcpyAsync(&flags, e->d_flags, sizeof flags, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  if (flags) return WF_E_OVERFLOW;
  float a = 0, b = 0, c = 0;
  (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
  (void)hipEventElapsedTime(&b, e->ev[1], e->ev[2]);
  if (want_psp) (void)hipEventElapsedTime(&c, e->ev[2], e->ev[3]);
  e->last.sieve_ms += a;
  e->last.test_ms += b;
  e->last.psp_ms += c;
  e->last.launches += 1;
  return WF_OK;
}

// Batched path: every pass of the batch in one HIP graph.  Pass k sieves into buffer set
// k & 1 on the engine's stream while the test kernels of pass k - 1 run on the second
// stream; a set is refilled only after the tests of the pass two back have finished.
static int run_batched(wf_engine* e, const std::vector<std::pair<std::vector<Piece>, int>>& passes,
                       const TestOut& to) {
  std::vector<PassPlan> plans(passes.size());
  std::vector<u64> soff(passes.size()), roff(passes.size()), foff(passes.size());
  u64 nsegs = 0, nruns = 0, nfills = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (size_t k = 0; k < passes.size(); k++) {
    plan_pass(e, passes[k].first, plans[k]);
    soff[k] = nsegs;
    roff[k] = nruns;
    foff[k] = nfills;
    nsegs += plans[k].segs.size();
    nruns += plans[k].fruns.size();
    nfills += (plans[k].fills.size() + 7) & ~(size_t)7;   // each pass's plan 32-byte aligned
  }
  e->last.plan_ms += ms_since(t0);
  if (nsegs == 0) return WF_OK;
  int rc;
  if ((rc = grow(e->d_segs, e->seg_cap, nsegs)) != WF_OK) return rc;
  if ((rc = grow(e->d_runs, e->run_cap, std::max<u64>(nruns, 1))) != WF_OK) return rc;
  if ((rc = grow(e->d_fills, e->fill_cap, std::max<u64>(nfills, 1))) != WF_OK) return rc;
  for (size_t k = 0; k < plans.size(); k++) {
    if (!plans[k].fills.empty())
      HIPCHK(hipMemcpyAsync(e->d_fills + foff[k], plans[k].fills.data(), plans[k].fills.size() * sizeof(u32),
                            hipMemcpyHostToDevice, e->st));
    if ((rc = size_set(e->bs[k & 1], plans[k])) != WF_OK) return rc;
    HIPCHK(hipMemcpyAsync(e->d_segs + soff[k], plans[k].segs.data(), plans[k].segs.size() * sizeof(SegDesc),
                          hipMemcpyHostToDevice, e->st));
    if (!plans[k].fruns.empty())
      HIPCHK(hipMemcpyAsync(e->d_runs + roff[k], plans[k].fruns.data(), plans[k].fruns.size() * sizeof(FillRun),
                            hipMemcpyHostToDevice, e->st));
  }
  HIPCHK(hipStreamSynchronize(e->st));
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeRelaxed));
  rc = WF_OK;
  if (hipEventRecord(e->gev[0], e->st) != hipSuccess || hipStreamWaitEvent(e->st2, e->gev[0], 0) != hipSuccess)
    rc = WF_E_DEVICE;
  for (size_t k = 0; k < plans.size() && rc == WF_OK; k++) {
    const PassPlan& P = plans[k];
    if (P.segs.empty()) continue;
    const int set = (int)(k & 1);
    hipEvent_t sieved = e->gev[1 + set], tested = e->gev[3 + set];
    if (k >= 2 && hipStreamWaitEvent(e->st, tested, 0) != hipSuccess) { rc = WF_E_DEVICE; break; }
    if ((rc = enqueue_sieve(e, P, e->d_segs + soff[k], e->d_runs + roff[k], e->d_fills + foff[k], e->bs[set], e->st,
                            to.acc)) != WF_OK)
      break;
    if (hipEventRecord(sieved, e->st) != hipSuccess || hipStreamWaitEvent(e->st2, sieved, 0) != hipSuccess) {
      rc = WF_E_DEVICE;
      break;
    }
    if ((rc = enqueue_test(e, P, e->d_segs + soff[k], e->bs[set], e->st2, passes[k].second, to, e->cfg.test_wgs)) !=
        WF_OK)
      break;
    if (hipEventRecord(tested, e->st2) != hipSuccess) { rc = WF_E_DEVICE; break; }
  }
  // join the test stream back into the capturing stream
  if (rc == WF_OK && (hipEventRecord(e->gev[0], e->st2) != hipSuccess ||
                      hipStreamWaitEvent(e->st, e->gev[0], 0) != hipSuccess))
    rc = WF_E_DEVICE;
  const hipError_t ce = hipStreamEndCapture(e->st, &graph);
  if (rc != WF_OK || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != WF_OK ? rc : WF_E_DEVICE;
  }
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return WF_E_DEVICE;
  }
  u32 flags = 0;
  rc = WF_OK;
  if (hipEventRecord(e->ev[0], e->st) != hipSuccess || hipGraphLaunch(exec, e->st) != hipSuccess ||
      hipEventRecord(e->ev[1], e->st) != hipSuccess ||
      hipMemcpyAsync(&flags, e->d_flags, sizeof flags, hipMemcpyDeviceToHost, e->st) != hipSuccess ||
      hipStreamSynchronize(e->st) != hipSuccess)
    rc = WF_E_DEVICE;
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (rc != WF_OK) return rc;
  if (flags) return WF_E_OVERFLOW;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
  e->last.total_ms += ms;
  e->last.launches += 1;
  return WF_OK;
}

// Events in (p, base, range) order: the same for any schedule of the kernels.
static bool cmp_dev_event(const DevEvent& a, const DevEvent& b) {
  if (a.p != b.p) return a.p < b.p;
  if (a.j != b.j) return a.j < b.j;
  if (a.rid != b.rid) return a.rid < b.rid;
  return a.type < b.type;
}

static void to_event(const DevEvent& d, u32 range, wf_event& o) {
  o.type = d.type;
  o.j = d.j;
  o.p = d.p;
  o.q = d.q;
  o.r_hi = d.r_hi;
  o.r_lo = d.r_lo;
  o.range = range;
  o.pad = 0;
}

// The p = 2 solutions of a range containing 2: one for every base b = 1 (mod 4), which
// scan_chunk records as a hit (wieferich.c:1143-1146); 2 itself is counted, never tested.
static bool has_two(u64 lo, u64 hi) { return lo <= 2 && (hi == 0 || hi > 2); }
static void push_two_events(const wf_engine* e, u32 rid, std::vector<DevEvent>& out) {
  for (int j = 0; j < e->nb; j++)
    if (e->bases[j] % 4 == 1) {
      DevEvent x;
      x.type = WF_EV_FOUND; x.rid = rid; x.j = (u32)j; x.pad = 0; x.p = 2; x.q = 0; x.r_hi = 0; x.r_lo = 1;
      out.push_back(x);
    }
}

// Scan n ranges; per-range totals into res[0..n), events into ev.
static int scan_ranges(wf_engine* e, const wf_range* r, u32 n, i64 near_T, wf_result* res, wf_event* ev,
                       u64 ev_cap, bool batched, bool want_psp = false, u64* psp = nullptr, u64 psp_cap = 0) {
  if (!e || !r || !res || n == 0 || (want_psp && n != 1)) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;   // the server thread owns the engine until wf_serve_stop
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
  const u64 ncks = (u64)n * e->nb;
  if ((rc = grow(e->d_cks, e->cks_cap, ncks)) != WF_OK) return rc;
  if ((rc = grow(e->d_sumq, e->sumq_cap, ncks)) != WF_OK) return rc;
  HIPCHK(hipMemsetAsync(e->d_acc, 0, n * sizeof(DevAcc), e->st));
  HIPCHK(hipMemsetAsync(e->d_cks, 0, ncks * sizeof(unsigned long long), e->st));
  HIPCHK(hipMemsetAsync(e->d_sumq, 0, ncks * sizeof(unsigned long long), e->st));
  HIPCHK(hipMemsetAsync(e->d_ctr, 0, 5 * sizeof(unsigned long long), e->st));
  if (want_psp && (rc = grow(e->d_psp, e->psp_cap, std::max<u64>(psp_cap, 1 << 16))) != WF_OK) return rc;
  HIPCHK(hipMemsetAsync(e->d_flags, 0, sizeof(u32), e->st));
  TestOut to;
  to.acc = e->d_acc;
  to.cks = e->d_cks;
  to.sumq = e->d_sumq;
  to.nb = (u32)e->nb;
  to.ev = e->d_ev;
  to.ev_cap = e->dev_ev_cap;
  to.nev = e->d_ctr + 1;
  to.near_T = near_T;
  to.nearv = nullptr;
  to.fault_p = e->cfg.fault_p;
  memset(&e->last, 0, sizeof e->last);
  // pieces by arithmetic class, then passes of at most cfg.chunk numbers each
  std::vector<Piece> all;
  for (u32 i = 0; i < n; i++) split_classes(e, r[i].lo, r[i].hi, i, all);
  const double chunk = (double)e->cfg.chunk;
  const u64 segspan = 30ull * e->cfg.seg_bytes;
  const u64 step = std::max<u64>(1, (u64)(chunk / (double)segspan)) * segspan;
  std::vector<std::pair<std::vector<Piece>, int>> passes;
  for (int cls = A_F64; cls < A_NCLS; cls++) {
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
          passes.push_back({batch, cls});
          batch.clear();
          in_batch = 0;
        }
        batch.push_back(part);
        in_batch += take;
        if (last) break;
        lo += step;
      }
    }
    if (!batch.empty()) passes.push_back({batch, cls});
  }
  if (batched) {
    if ((rc = run_batched(e, passes, to)) != WF_OK) return rc;
  } else {
    for (auto& ps : passes)
      if ((rc = run_pieces(e, ps.first, ps.second, to, want_psp)) != WF_OK) return rc;
  }
  // the strong Lucas test of the pseudoprime list (one range: want_psp comes from wf_scan)
  if (want_psp) {
    hipLaunchKernelGGL(k_psp_lucas, dim3(64), dim3(256), 0, e->st, e->d_psp, e->d_ctr + 2, e->psp_cap, e->d_acc);
    HIPCHK(hipGetLastError());
  }
  // results
  std::vector<DevAcc> acc(n);
  std::vector<unsigned long long> hcks(ncks), hsumq(ncks);
  unsigned long long nev = 0;
  HIPCHK(hipMemcpyAsync(acc.data(), e->d_acc, n * sizeof(DevAcc), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(hcks.data(), e->d_cks, ncks * sizeof(unsigned long long), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(hsumq.data(), e->d_sumq, ncks * sizeof(unsigned long long), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(&nev, e->d_ctr + 1, sizeof nev, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  const u64 got = std::min<u64>(nev, e->dev_ev_cap);
  std::vector<DevEvent> dev(got);
  if (got) HIPCHK(hipMemcpy(dev.data(), e->d_ev, got * sizeof(DevEvent), hipMemcpyDeviceToHost));
  for (u32 i = 0; i < n; i++)
    if (has_two(r[i].lo, r[i].hi)) {
      const size_t before = dev.size();
      push_two_events(e, i, dev);
      acc[i].nev += dev.size() - before;
    }
  std::sort(dev.begin(), dev.end(), cmp_dev_event);
  const u64 ndev = dev.size();
  u64 tested = 0;
  for (u32 i = 0; i < n; i++) {
    wf_result& o = res[i];
    memset(&o, 0, sizeof o);
    o.primes = acc[i].primes;
    // the prime 2 is counted, never tested (wieferich.c:1142-1146)
    if (has_two(r[i].lo, r[i].hi)) o.primes++;
    o.tests = acc[i].tests;
    o.near = acc[i].near;
    o.flt = acc[i].flt;
    for (int j = 0; j < e->nb; j++) {
      o.cks[j] = hcks[(u64)i * e->nb + j];
      o.sumq[j] = hsumq[(u64)i * e->nb + j];
    }
    o.nev = acc[i].nev;
    tested += o.tests;
  }
  for (u64 k = 0; k < ndev && k < ev_cap; k++) to_event(dev[k], dev[k].rid, ev[k]);
  e->last.tested = tested;
  if (!batched) e->last.total_ms = e->last.sieve_ms + e->last.test_ms;
  for (u32 i = 0; i < n; i++) {
    res[i].npsp = acc[i].npsp;
    res[i].nbpsw = acc[i].nbpsw;
  }
  unsigned long long proved = 0, ptests = 0;
  HIPCHK(hipMemcpy(&proved, e->d_ctr + 3, sizeof proved, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&ptests, e->d_ctr + 4, sizeof ptests, hipMemcpyDeviceToHost));
  e->last.proved = proved;
  e->last.psp_tests = ptests;
  if (want_psp) {
    unsigned long long np = 0;
    HIPCHK(hipMemcpy(&np, e->d_ctr + 2, sizeof np, hipMemcpyDeviceToHost));
    std::vector<u64> lst(std::min<u64>(np, e->psp_cap));
    if (!lst.empty()) HIPCHK(hipMemcpy(lst.data(), e->d_psp, lst.size() * sizeof(u64), hipMemcpyDeviceToHost));
    std::sort(lst.begin(), lst.end());
    for (u64 k = 0; k < lst.size() && k < psp_cap; k++) psp[k] = lst[k];
    if (np > e->psp_cap) return WF_E_OVERFLOW;
  }
  if (nev > e->dev_ev_cap) return WF_E_OVERFLOW;
  return WF_OK;
}

extern "C" int wf_scan(wf_engine* e, u64 lo, u64 hi, int64_t near_T, wf_result* out, wf_event* ev, u64 ev_cap,
                       int want_psp, u64* psp, u64 psp_cap) {
  if (!e || !out || (ev_cap && !ev) || (want_psp && psp_cap && !psp)) return WF_E_ARG;
  wf_range r = {lo, hi};
  return scan_ranges(e, &r, 1, near_T, out, ev, ev_cap, false, want_psp != 0, psp, psp_cap);
}

extern "C" int wf_scan_batch(wf_engine* e, const wf_range* r, u32 n, int64_t near_T, wf_result* per_range,
                             wf_event* ev, u64 ev_cap) {
  if (!e || !r || !per_range || (ev_cap && !ev)) return WF_E_ARG;
  return scan_ranges(e, r, n, near_T, per_range, ev, ev_cap, e->cfg.pipeline == WF_PIPELINE_GRAPH);
}

static int quotient_args(wf_engine* e, uint32_t b, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n) {
  if (!e || !p || !q || !r_hi || !r_lo || b < 2) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  for (u64 i = 0; i < n; i++)
    if (p[i] < 3 || !(p[i] & 1)) return WF_E_ARG;
  return WF_OK;
}

// k_quot over n values staged in d[0 .. 4n): p in, then q, r_hi, r_lo out.
static int quotients_in(wf_engine* e, uint32_t b, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n, u64* d) {
  int rc;
  if ((rc = put_base(WF_MAXB, b)) != WF_OK) return rc;
  HIPCHK(hipMemcpyAsync(d, p, n * sizeof(u64), hipMemcpyHostToDevice, e->st));
  hipLaunchKernelGGL(k_quot, dim3((u32)((n + 255) / 256)), dim3(256), 0, e->st, d, d + n, d + 2 * n, d + 3 * n, n,
                     e->cfg.arith == 1 ? 1 : 0, (u32)WF_MAXB);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(q, d + n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(r_hi, d + 2 * n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(r_lo, d + 3 * n, n * sizeof(u64), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  return WF_OK;
}

extern "C" int wf_quotients_base(wf_engine* e, uint32_t b, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n) {
  int rc;
  if ((rc = quotient_args(e, b, p, q, r_hi, r_lo, n)) != WF_OK || n == 0) return rc;
  HIPCHK(hipSetDevice(e->dev));
  u64* d = nullptr;
  if (hipMalloc(&d, 4 * n * sizeof(u64)) != hipSuccess) return WF_E_NOMEM;
  rc = quotients_in(e, b, p, q, r_hi, r_lo, n, d);
  (void)hipStreamSynchronize(e->st);    // an early error return leaves nothing queued on d
  (void)hipFree(d);
  return rc;
}

// The no-allocation variant: the staging area is the device's static scratch array, whose
// address wf_open took once per device (scratch_addr).  Every call on a GPU shares it, so
// calls that overlap on one GPU race; the area holds the last call's values afterwards.
extern "C" int wf_quotients_static(wf_engine* e, uint32_t b, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n) {
  int rc;
  if ((rc = quotient_args(e, b, p, q, r_hi, r_lo, n)) != WF_OK || n == 0) return rc;
  if (n > WF_STATIC_QUOT_MAX) return WF_E_LIMIT;
  if (!e->scratch) return WF_E_DEVICE;
  HIPCHK(hipSetDevice(e->dev));
  return quotients_in(e, b, p, q, r_hi, r_lo, n, e->scratch);
}

extern "C" int wf_quotients(wf_engine* e, const u64* p, u64* q, u64* r_hi, u64* r_lo, u64 n) {
  return wf_quotients_base(e, 2, p, q, r_hi, r_lo, n);
}

extern "C" int wf_bpsw(wf_engine* e, const u64* n, uint8_t* flags, u64 count) {
  if (!e || (count && (!n || !flags))) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  if (count == 0) return WF_OK;
  HIPCHK(hipSetDevice(e->dev));
  u64* d = nullptr;
  if (hipMalloc(&d, count * (sizeof(u64) + 1)) != hipSuccess) return WF_E_NOMEM;
  unsigned char* dout = (unsigned char*)(d + count);
  int rc = WF_OK;
  if (hipMemcpyAsync(d, n, count * sizeof(u64), hipMemcpyHostToDevice, e->st) != hipSuccess) rc = WF_E_DEVICE;
  if (rc == WF_OK) {
    hipLaunchKernelGGL(k_bpsw, dim3((u32)((count + 255) / 256)), dim3(256), 0, e->st, d, dout, count);
    if (hipGetLastError() != hipSuccess) rc = WF_E_DEVICE;
  }
  if (rc == WF_OK && (hipMemcpyAsync(flags, dout, count, hipMemcpyDeviceToHost, e->st) != hipSuccess ||
                      hipStreamSynchronize(e->st) != hipSuccess))
    rc = WF_E_DEVICE;
  (void)hipFree(d);
  return rc;
}

extern "C" int wf_sieve_stats(wf_engine* e, u64 lo, u64 hi, wf_stats* out) {
  if (!e || !out || hi <= lo || hi > WF_STATS_MAXHI) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  HIPCHK(hipSetDevice(e->dev));
  int rc;
  if ((rc = ensure_sieving_primes(e, isqrt_u64(hi - 1 + STATS_OV) + 1)) != WF_OK) return rc;
  if ((rc = grow(e->d_acc, e->acc_cap, 1)) != WF_OK) return rc;
  if (!e->d_stats && hipMalloc(&e->d_stats, sizeof(StatsOut)) != hipSuccess) return WF_E_NOMEM;
  HIPCHK(hipMemsetAsync(e->d_stats, 0, offsetof(StatsOut, first), e->st));
  HIPCHK(hipMemsetAsync(e->d_stats->first, 0xFF, sizeof e->d_stats->first, e->st));
  HIPCHK(hipMemsetAsync(e->d_flags, 0, sizeof(u32), e->st));
  HIPCHK(hipMemsetAsync(e->d_ctr + 3, 0, sizeof(unsigned long long), e->st));
  memset(out, 0, sizeof *out);
  const u64 segspan = 30ull * e->cfg.seg_bytes;
  // passes of at most cfg.chunk numbers (and < 2^46, for the gap key's offset field)
  const u64 step = std::max<u64>(1, std::min<u64>(e->cfg.chunk, 1ull << 46) / segspan) * segspan;
  const int cus = e->prop.multiProcessorCount;
  for (u64 a = lo; a < hi;) {
    const u64 b = hi - a <= step ? hi : a + step;
    std::vector<Piece> pieces = {{a, b + STATS_OV, 0, A_F64}};
    PassPlan P;
    plan_pass(e, pieces, P);
    if ((rc = size_set(e->bs[0], P)) != WF_OK) return rc;
    if ((rc = upload_pass(e, P)) != WF_OK) return rc;
    if ((rc = enqueue_sieve(e, P, e->d_segs, e->d_runs, e->d_fills, e->bs[0], e->st, e->d_acc)) != WF_OK) return rc;
    HIPCHK(hipMemsetAsync(&e->d_stats->gapkey, 0, sizeof(unsigned long long), e->st));
    const u32 tps = (P.pcap + 255) / 256;
    const u64 nseg = P.segs.size();
    hipLaunchKernelGGL(k_stats, dim3((u32)std::min<u64>(nseg * tps, (u64)cus * 16)), dim3(256), 0, e->st, e->d_segs,
                       (u32)nseg, e->bs[0].off, e->bs[0].cnt, P.pcap, tps, a, b, e->d_stats, e->d_flags);
    HIPCHK(hipGetLastError());
    unsigned long long gk = 0;
    u32 flags = 0;
    HIPCHK(hipMemcpyAsync(&gk, &e->d_stats->gapkey, sizeof gk, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipMemcpyAsync(&flags, e->d_flags, sizeof flags, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));   // also keeps P's host arrays alive for the uploads
    if (flags) return WF_E_OVERFLOW;
    const u64 gap = gk >> 48, gp = a + (((1ull << 48) - 1) - (gk & ((1ull << 48) - 1)));
    if (gk && gap > out->max_gap) {       // ties keep the earlier pass's smaller p
      out->max_gap = gap;
      out->max_gap_p = gp;
    }
    a = b;
  }
  StatsOut h;
  HIPCHK(hipMemcpy(&h, e->d_stats, sizeof h, hipMemcpyDeviceToHost));
  out->primes = h.primes;
  out->twins = h.twins;
  for (int n = 0; n < 64; n++) {
    out->runs[n] = h.runs[n];
    out->first[n] = h.first[n] == ~0ull ? 0 : h.first[n];
  }
  // the prime 2 is not on the wheel: L(2) = 2 (gaps 1 < 2, then 2), its gap is 1
  if (lo <= 2 && hi > 2) {
    out->primes++;
    out->runs[2]++;
    out->first[2] = 2;
    if (out->max_gap < 1) {
      out->max_gap = 1;
      out->max_gap_p = 2;
    }
  }
  return WF_OK;
}

extern "C" int wf_primes(wf_engine* e, u64 lo, u64 hi, u64* out, u64 cap, u64* count) {
  if (!e || !count || (cap && !out) || (hi != 0 && hi <= lo)) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  HIPCHK(hipSetDevice(e->dev));
  const u64 top = hi ? hi - 1 : 0xFFFFFFFFFFFFFFFFull;
  int rc;
  if ((rc = ensure_sieving_primes(e, std::max<u64>(isqrt_u64(top) + 1, 64))) != WF_OK) return rc;
  if ((rc = grow(e->d_acc, e->acc_cap, 1)) != WF_OK) return rc;
  HIPCHK(hipMemsetAsync(e->d_acc, 0, sizeof(DevAcc), e->st));
  HIPCHK(hipMemsetAsync(e->d_flags, 0, sizeof(u32), e->st));
  std::vector<u64> all;
  if (lo <= 2 && (hi == 0 || hi > 2)) all.push_back(2);
  // passes of at most cfg.chunk numbers, sieve only
  const u64 segspan = 30ull * e->cfg.seg_bytes;
  const u64 step = std::max<u64>(1, e->cfg.chunk / segspan) * segspan;
  u64 cur = lo;
  for (;;) {
    const bool last = span_d(cur, hi) <= (double)step;
    std::vector<Piece> pieces = {{cur, last ? hi : cur + step, 0, A_F64}};
    PassPlan P;
    plan_pass(e, pieces, P);
    if ((rc = size_set(e->bs[0], P)) != WF_OK) return rc;
    if ((rc = upload_pass(e, P)) != WF_OK) return rc;
    if ((rc = enqueue_sieve(e, P, e->d_segs, e->d_runs, e->d_fills, e->bs[0], e->st, e->d_acc)) != WF_OK) return rc;
    std::vector<u32> cnt(P.segs.size());
    u32 flags = 0;
    HIPCHK(hipMemcpyAsync(cnt.data(), e->bs[0].cnt, cnt.size() * sizeof(u32), hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipMemcpyAsync(&flags, e->d_flags, sizeof flags, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    if (flags) return WF_E_OVERFLOW;
    std::vector<u32> offs;
    for (size_t s = 0; s < P.segs.size(); s++) {
      offs.resize(cnt[s]);
      if (cnt[s]) HIPCHK(hipMemcpy(offs.data(), e->bs[0].off + s * P.pcap, cnt[s] * sizeof(u32), hipMemcpyDeviceToHost));
      for (u32 k = 0; k < cnt[s]; k++) all.push_back(P.segs[s].base + offs[k]);
    }
    if (last) break;
    cur += step;
  }
  std::sort(all.begin(), all.end());
  *count = all.size();
  for (u64 i = 0; i < all.size() && i < cap; i++) out[i] = all[i];
  return WF_OK;
}

// Switch the engine to configuration c (segment size, sieve workgroup, bucket threshold).
static void apply_config(wf_engine* e, const wf_config& c) {
  e->cfg.seg_bytes = c.seg_bytes;
  e->cfg.sieve_block = c.sieve_block;
  e->cfg.bucket_min = c.bucket_min;
  e->sbshift = __builtin_ctz(e->cfg.seg_bytes);
  if (!e->h_sp.empty()) set_bucket_min(e);
  e->slices.clear();
}

// Autotuner: time the sieve stage (bucket fill + segment sieve + extraction) of a sample
// of [lo, hi) for every segment size 16-128 KiB, sieve workgroup of 4-16 wavefronts and
// bucket threshold 2^15-2^20, keep the fastest.  The test kernel does not depend on them.
// Time the sieve stage of the given pieces for every configuration, keep the fastest.
static int autotune_pieces(wf_engine* e, const std::vector<Piece>& pieces, double numbers, wf_config* best,
                           double* best_ms) {
  int rc;
  if ((rc = grow(e->d_acc, e->acc_cap, 1)) != WF_OK) return rc;
  const wf_config orig = e->cfg;
  wf_config bestc = orig;
  double bestt = 1e300;
  const u32 segs[] = {16384, 32768, 65536, 131072}, blocks[] = {256, 512, 1024},
            bmins[] = {32768, 65536, 131072, 262144, 524288, 1048576};
  for (u32 sb : segs)
    for (u32 bl : blocks)
      for (u32 bm : bmins) {
        if (sb + 4096 > e->prop.sharedMemPerBlock) continue;
        wf_config c = orig;
        c.seg_bytes = sb;
        c.sieve_block = bl;
        c.bucket_min = bm;
        apply_config(e, c);
        PassPlan P;
        plan_pass(e, pieces, P);
        if ((rc = size_set(e->bs[0], P)) != WF_OK) { apply_config(e, orig); return rc; }
        if ((rc = upload_pass(e, P)) != WF_OK) { apply_config(e, orig); return rc; }
        HIPCHK(hipMemsetAsync(e->d_flags, 0, sizeof(u32), e->st));
        float ms = 0;
        for (int rep = 0; rep < 2; rep++) {   // the first run warms caches and code
          HIPCHK(hipEventRecord(e->ev[0], e->st));
          if ((rc = enqueue_sieve(e, P, e->d_segs, e->d_runs, e->d_fills, e->bs[0], e->st, e->d_acc)) != WF_OK) {
            apply_config(e, orig);
            return rc;
          }
          HIPCHK(hipEventRecord(e->ev[1], e->st));
          HIPCHK(hipStreamSynchronize(e->st));
          (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
        }
        if (ms < bestt) {
          bestt = ms;
          bestc = c;
        }
      }
  apply_config(e, bestc);
  if (best) *best = e->cfg;
  if (best_ms) *best_ms = bestt / (numbers / 1e9);   // ms per 1e9 numbers
  return WF_OK;
}

extern "C" int wf_autotune(wf_engine* e, u64 lo, u64 hi, wf_config* best, double* best_ms) {
  if (!e || (hi != 0 && hi <= lo)) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  HIPCHK(hipSetDevice(e->dev));
  const u64 top = hi ? hi - 1 : 0xFFFFFFFFFFFFFFFFull;
  int rc;
  if ((rc = ensure_sieving_primes(e, std::max<u64>(isqrt_u64(top) + 1, 64))) != WF_OK) return rc;
  // sample: up to 3e9 numbers from the start of the range
  const u64 sample = (u64)std::min(span_d(lo, hi), 3.0e9);
  return autotune_pieces(e, {{lo, lo + sample, 0, A_F64}}, (double)sample, best, best_ms);
}

extern "C" int wf_autotune_batch(wf_engine* e, const wf_range* r, u32 n, wf_config* best, double* best_ms) {
  if (!e || !r || n == 0) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  HIPCHK(hipSetDevice(e->dev));
  // sample: the ranges in order, up to 3e9 numbers (the last one cut), packed as one pass
  std::vector<Piece> pieces;
  double total = 0;
  u64 maxtop = 0;
  for (u32 i = 0; i < n && total < 3.0e9; i++) {
    if (r[i].hi != 0 && r[i].hi <= r[i].lo) return WF_E_ARG;
    const double len = std::min(span_d(r[i].lo, r[i].hi), 3.0e9 - total);
    const u64 hi = r[i].lo + (u64)len;
    pieces.push_back({r[i].lo, hi, 0, A_F64});
    maxtop = std::max<u64>(maxtop, hi ? hi - 1 : 0xFFFFFFFFFFFFFFFFull);
    total += len;
  }
  int rc;
  if ((rc = ensure_sieving_primes(e, std::max<u64>(isqrt_u64(maxtop) + 1, 64))) != WF_OK) return rc;
  return autotune_pieces(e, pieces, total, best, best_ms);
}

extern "C" int wf_last_timing(wf_engine* e, wf_timing* t) {
  if (!e || !t) return WF_E_ARG;
  if (e->srv) return WF_E_BUSY;
  *t = e->last;
  return WF_OK;
}

// ---------------------------------------------------------------------------------
// serving: continuous batching of range queries
// ---------------------------------------------------------------------------------
//
// An admitted request owns a slot: a row of accumulators, checksums and its near-hit bound in
// HBM.  The server thread runs steps.  A step takes the next piece of every request that still
// has numbers to schedule (at least min_piece numbers, at most an even share of the step
// budget), one pass per arithmetic class, and sieves and tests all pieces of a pass in the
// same launches (range id = slot).  Two steps are in flight at once: while step k runs, the
// thread plans and enqueues step k + 1, which takes the requests that arrived meanwhile.
// Passes alternate between the engine's two buffer sets as in the batched path: a pass
// sieves on the engine's stream while the previous pass's tests run on the second stream.
// Each step has its own event buffer and copies the slots' totals to pinned host memory after
// its last test; the thread then routes the step's events to their requests by slot, and a
// request whose last piece was in the step is finished.  A slot handed to a new request is
// cleared by a kernel on the engine's stream before the step that first uses it, so no
// memset runs while steps are in flight.

#define SERVE_EVCAP (1u << 18)      // device events per step
#define SERVE_INIT_MAX 32           // slots set up per k_slot_init launch

struct SlotInit {
  u32 n, pad;
  u32 slot[SERVE_INIT_MAX];
  i64 near_T[SERVE_INIT_MAX];
};

// Clear the accumulators of newly admitted slots and set their near-hit bounds.
__global__ void k_slot_init(DevAcc* __restrict__ acc, unsigned long long* __restrict__ cks,
                            unsigned long long* __restrict__ sumq, i64* __restrict__ nearv, u32 nb, SlotInit s) {
  for (u32 k = threadIdx.x; k < s.n; k += blockDim.x) {
    const u32 sl = s.slot[k];
    DevAcc z = {0, 0, 0, 0, 0, 0, 0};
    acc[sl] = z;
    nearv[sl] = s.near_T[k];
    for (u32 j = 0; j < nb; j++) {
      cks[(u64)sl * nb + j] = 0;
      sumq[(u64)sl * nb + j] = 0;
    }
  }
}

struct ServeReq {
  u64 lo, hi, next;        // hi == 0: 2^64; next = first number not yet scheduled
  i64 near_T;
  u32 slot;
  bool scheduled;          // every number is in a step
  bool done;
  int rc;                  // WF_OK, or the error the request ended with
  wf_result res;
  std::vector<DevEvent> ev;   // its events so far
};

struct ServeStep {         // a step in flight
  std::vector<PassPlan> plans;   // kept until the step is done
  std::vector<int> arith;
  std::vector<u64> finishing;    // tickets whose last piece is in this step
  u32 nslot = 0;                 // slots copied back (highest slot in use + 1)
  SegDesc* d_segs = nullptr;
  u64 seg_cap = 0;
  FillRun* d_runs = nullptr;
  u64 run_cap = 0;
  u32* d_fills = nullptr;        // fill plans (FillGroup words and slice tables)
  u64 fill_cap = 0;
  DevEvent* d_ev = nullptr;
  unsigned long long* d_nev = nullptr;
  DevAcc* h_acc = nullptr;       // pinned
  unsigned long long* h_cks = nullptr;
  unsigned long long* h_sumq = nullptr;
  unsigned long long* h_nev = nullptr;
  u32* h_flags = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t tev[A_NCLS][4] = {};   // per pass (one per class): sieve start / end, test start / end
};

struct Server {
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  int err = WF_OK;
  u32 slots = 0, active = 0, nb = 1;
  u64 step_numbers = 0, min_piece = 0;
  std::vector<u64> slot_ticket;    // 0 = free
  std::vector<std::pair<u32, i64>> fresh;   // slots admitted since the last step, with near_T
  std::unordered_map<u64, ServeReq> reqs;
  u64 next_ticket = 1;
  DevAcc* d_acc = nullptr;         // per slot
  unsigned long long* d_cks = nullptr;    // per slot and base
  unsigned long long* d_sumq = nullptr;
  i64* d_nearv = nullptr;          // per slot
  ServeStep step[2];
  std::vector<int> inflight;       // step contexts in flight, oldest first
  int next_ctx = 0;
  u64 gpass = 0;                   // passes enqueued so far (their buffer sets alternate)
  u64 planned = 0;                 // steps planned
  u64 steps = 0;                   // steps finished
  wf_timing tim = {};              // kernel times and tests of the session (wf_last_timing after stop)
};

// Wait for the oldest step in flight and hand its results to the requests.
static int serve_finish_oldest(wf_engine* e, Server* S) {
  ServeStep& c = S->step[S->inflight.front()];
  S->inflight.erase(S->inflight.begin());
  if (hipEventSynchronize(c.done) != hipSuccess) return WF_E_DEVICE;
  if (*c.h_flags) return WF_E_OVERFLOW;   // a sieve bound was exceeded (bug guard)
  for (size_t k = 0; k < c.plans.size(); k++) {
    float a = 0, b = 0;
    (void)hipEventElapsedTime(&a, c.tev[k][0], c.tev[k][1]);
    (void)hipEventElapsedTime(&b, c.tev[k][2], c.tev[k][3]);
    S->tim.sieve_ms += a;
    S->tim.test_ms += b;
    S->tim.launches++;
  }
  const u64 nev = *c.h_nev;
  std::vector<DevEvent> ev(std::min<u64>(nev, SERVE_EVCAP));
  if (!ev.empty() && hipMemcpy(ev.data(), c.d_ev, ev.size() * sizeof(DevEvent), hipMemcpyDeviceToHost) != hipSuccess)
    return WF_E_DEVICE;
  std::lock_guard<std::mutex> lk(S->mu);
  S->steps++;
  for (const DevEvent& x : ev) {
    const u64 t = x.rid < S->slots ? S->slot_ticket[x.rid] : 0;
    auto it = S->reqs.find(t);
    if (it != S->reqs.end()) it->second.ev.push_back(x);
  }
  for (u64 t : c.finishing) {
    auto it = S->reqs.find(t);
    if (it == S->reqs.end()) continue;
    ServeReq& q = it->second;
    const u32 k = q.slot;
    const DevAcc& a = c.h_acc[k];
    memset(&q.res, 0, sizeof q.res);
    q.res.primes = a.primes;
    q.res.tests = a.tests;
    q.res.near = a.near;
    q.res.flt = a.flt;
    for (u32 j = 0; j < S->nb; j++) {
      q.res.cks[j] = c.h_cks[(size_t)k * S->nb + j];
      q.res.sumq[j] = c.h_sumq[(size_t)k * S->nb + j];
    }
    q.res.nev = a.nev;
    // every event the kernels counted must have reached the step's buffer
    q.rc = q.ev.size() == a.nev ? WF_OK : WF_E_OVERFLOW;
    if (has_two(q.lo, q.hi)) {       // 2 is counted, never tested; its solutions as in wf_scan
      q.res.primes++;
      const size_t before = q.ev.size();
      push_two_events(e, k, q.ev);
      q.res.nev += q.ev.size() - before;
    }
    std::sort(q.ev.begin(), q.ev.end(), cmp_dev_event);
    q.done = true;
    S->tim.tested += q.res.tests;
    S->slot_ticket[k] = 0;
    S->active--;
  }
  S->cv.notify_all();
  return WF_OK;
}

static int serve_drain(wf_engine* e, Server* S) {
  int rc = WF_OK;
  while (!S->inflight.empty() && rc == WF_OK) rc = serve_finish_oldest(e, S);
  return rc;
}

// A failed step ends every request with its error; the server then takes no more.
static void serve_fail(wf_engine* e, Server* S, int rc) {
  (void)hipStreamSynchronize(e->st);
  (void)hipStreamSynchronize(e->st2);
  S->inflight.clear();
  std::lock_guard<std::mutex> lk(S->mu);
  S->err = rc;
  for (auto& kv : S->reqs)
    if (!kv.second.done) {
      kv.second.done = true;
      kv.second.rc = rc;
    }
  std::fill(S->slot_ticket.begin(), S->slot_ticket.end(), 0);
  S->fresh.clear();
  S->active = 0;
  S->cv.notify_all();
}

static bool set_fits(const BufSet& B, const PassPlan& P) {
  const u64 nseg = P.segs.size();
  return B.cnt_cap >= nseg && B.off_cap >= nseg * P.pcap && (!P.bcap || (B.bcnt_cap >= nseg &&
         B.bdata_cap >= nseg * (u64)P.bcap)) && (!P.nprove || B.tcnt_cap >= nseg * (u64)prove_tiles(P));
}

// Plan the passes of one step's pieces and enqueue them into context c.
static int serve_enqueue(wf_engine* e, Server* S, int ci, const std::vector<Piece>& all,
                         const std::vector<std::pair<u32, i64>>& inits) {
  ServeStep& c = S->step[ci];
  int rc;
  u64 maxtop = 0;
  for (const Piece& p : all) maxtop = std::max<u64>(maxtop, p.hi ? p.hi - 1 : 0xFFFFFFFFFFFFFFFFull);
  const u64 need = std::min<u64>(std::max<u64>(isqrt_u64(maxtop) + 1, 64), 0xFFFFFFFFull);
  if (need > e->sp_lim) {            // the table is rebuilt: nothing in flight may read it
    if ((rc = serve_drain(e, S)) != WF_OK) return rc;
    if ((rc = ensure_sieving_primes(e, need)) != WF_OK) return rc;
  }
  c.plans.clear();
  c.arith.clear();
  for (int cls = A_F64; cls < A_NCLS; cls++) {
    std::vector<Piece> ps;
    for (const Piece& p : all)
      if (p.arith == cls) ps.push_back(p);
    if (ps.empty()) continue;
    c.plans.emplace_back();
    plan_pass(e, ps, c.plans.back());
    c.arith.push_back(cls);
  }
  u64 nsegs = 0, nruns = 0, nfills = 0;
  for (const PassPlan& P : c.plans) {
    nsegs += P.segs.size();
    nruns += P.fruns.size();
    nfills += (P.fills.size() + 7) & ~(size_t)7;   // each pass's plan 32-byte aligned
  }
  if ((rc = grow(c.d_segs, c.seg_cap, std::max<u64>(nsegs, 1))) != WF_OK) return rc;
  if ((rc = grow(c.d_runs, c.run_cap, std::max<u64>(nruns, 1))) != WF_OK) return rc;
  if ((rc = grow(c.d_fills, c.fill_cap, std::max<u64>(nfills, 1))) != WF_OK) return rc;
  // buffer sets: a pass that needs more room than its set has waits for everything in flight
  for (size_t k = 0; k < c.plans.size(); k++) {
    BufSet& B = e->bs[(S->gpass + k) & 1];
    if (!set_fits(B, c.plans[k])) {
      if ((rc = serve_drain(e, S)) != WF_OK) return rc;
      if ((rc = size_set(B, c.plans[k])) != WF_OK) return rc;
    }
  }
  // new slots, the step's event counter, then the descriptors
  for (size_t k = 0; k < inits.size(); k += SERVE_INIT_MAX) {
    SlotInit si;
    memset(&si, 0, sizeof si);
    si.n = (u32)std::min<size_t>(SERVE_INIT_MAX, inits.size() - k);
    for (u32 i = 0; i < si.n; i++) {
      si.slot[i] = inits[k + i].first;
      si.near_T[i] = inits[k + i].second;
    }
    hipLaunchKernelGGL(k_slot_init, dim3(1), dim3(64), 0, e->st, S->d_acc, S->d_cks, S->d_sumq, S->d_nearv, S->nb, si);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, e->st, (u32*)c.d_nev, (u64)2);
  HIPCHK(hipGetLastError());
  u64 so = 0, ro = 0, fo = 0;
  std::vector<u64> soff, roff, foff;
  for (const PassPlan& P : c.plans) {
    soff.push_back(so);
    roff.push_back(ro);
    foff.push_back(fo);
    HIPCHK(hipMemcpyAsync(c.d_segs + so, P.segs.data(), P.segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, e->st));
    if (!P.fruns.empty())
      HIPCHK(hipMemcpyAsync(c.d_runs + ro, P.fruns.data(), P.fruns.size() * sizeof(FillRun), hipMemcpyHostToDevice, e->st));
    if (!P.fills.empty())
      HIPCHK(hipMemcpyAsync(c.d_fills + fo, P.fills.data(), P.fills.size() * sizeof(u32), hipMemcpyHostToDevice,
                            e->st));
    so += P.segs.size();
    ro += P.fruns.size();
    fo += (P.fills.size() + 7) & ~(size_t)7;
  }
  TestOut to;
  to.acc = S->d_acc;
  to.cks = S->d_cks;
  to.sumq = S->d_sumq;
  to.nb = S->nb;
  to.ev = c.d_ev;
  to.ev_cap = SERVE_EVCAP;
  to.nev = c.d_nev;
  to.near_T = 0;
  to.nearv = S->d_nearv;
  to.fault_p = e->cfg.fault_p;
  for (size_t k = 0; k < c.plans.size(); k++) {
    const PassPlan& P = c.plans[k];
    const int set = (int)((S->gpass + k) & 1);
    hipEvent_t sieved = e->gev[1 + set], tested = e->gev[3 + set];
    if (S->gpass + k >= 2) HIPCHK(hipStreamWaitEvent(e->st, tested, 0));
    HIPCHK(hipEventRecord(c.tev[k][0], e->st));
    if ((rc = enqueue_sieve(e, P, c.d_segs + soff[k], c.d_runs + roff[k], c.d_fills + foff[k], e->bs[set], e->st,
                            S->d_acc)) != WF_OK)
      return rc;
    HIPCHK(hipEventRecord(c.tev[k][1], e->st));
    HIPCHK(hipEventRecord(sieved, e->st));
    HIPCHK(hipStreamWaitEvent(e->st2, sieved, 0));
    HIPCHK(hipEventRecord(c.tev[k][2], e->st2));
    if ((rc = enqueue_test(e, P, c.d_segs + soff[k], e->bs[set], e->st2, c.arith[k], to)) != WF_OK) return rc;
    HIPCHK(hipEventRecord(c.tev[k][3], e->st2));
    HIPCHK(hipEventRecord(tested, e->st2));
  }
  S->gpass += c.plans.size();
  // the slots' totals after this step (a request finishing here has no piece in a later one)
  const size_t nck = (size_t)c.nslot * S->nb * sizeof(unsigned long long);
  HIPCHK(hipMemcpyAsync(c.h_acc, S->d_acc, (size_t)c.nslot * sizeof(DevAcc), hipMemcpyDeviceToHost, e->st2));
  HIPCHK(hipMemcpyAsync(c.h_cks, S->d_cks, nck, hipMemcpyDeviceToHost, e->st2));
  HIPCHK(hipMemcpyAsync(c.h_sumq, S->d_sumq, nck, hipMemcpyDeviceToHost, e->st2));
  HIPCHK(hipMemcpyAsync(c.h_nev, c.d_nev, sizeof(unsigned long long), hipMemcpyDeviceToHost, e->st2));
  HIPCHK(hipMemcpyAsync(c.h_flags, e->d_flags, sizeof(u32), hipMemcpyDeviceToHost, e->st2));
  HIPCHK(hipEventRecord(c.done, e->st2));
  S->inflight.push_back(ci);
  return WF_OK;
}

static void serve_loop(wf_engine* e) {
  Server* S = e->srv;
  (void)hipSetDevice(e->dev);
  for (;;) {
    std::vector<Piece> all;
    std::vector<std::pair<u32, i64>> inits;
    int ci = -1;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto schedulable = [&]() {
        for (u32 k = 0; k < S->slots; k++) {
          const u64 t = S->slot_ticket[k];
          if (t && !S->reqs[t].scheduled) return true;
        }
        return false;
      };
      S->cv.wait(lk, [&] { return S->stop || !S->inflight.empty() || (S->err == WF_OK && schedulable()); });
      if (S->stop) {
        lk.unlock();
        (void)hipStreamSynchronize(e->st);
        (void)hipStreamSynchronize(e->st2);
        S->inflight.clear();
        return;
      }
      if (S->err == WF_OK && S->inflight.size() < 2 && schedulable()) {
        ci = S->next_ctx;
        ServeStep& c = S->step[ci];
        c.finishing.clear();
        c.nslot = 0;
        u32 nsched = 0;
        for (u32 k = 0; k < S->slots; k++)
          if (S->slot_ticket[k] && !S->reqs[S->slot_ticket[k]].scheduled) nsched++;
        const u64 share = std::max<u64>(S->min_piece, S->step_numbers / std::max<u32>(nsched, 1));
        for (u32 k = 0; k < S->slots; k++) {
          const u64 t = S->slot_ticket[k];
          if (!t) continue;
          c.nslot = k + 1;
          ServeReq& q = S->reqs[t];
          if (q.scheduled) continue;
          const bool last = span_d(q.next, q.hi) <= (double)share;
          const u64 end = last ? q.hi : q.next + share;
          split_classes(e, q.next, end, k, all);
          q.next = end;
          if (last) {
            q.scheduled = true;
            c.finishing.push_back(t);
          }
        }
        inits.swap(S->fresh);
      }
    }
    int rc;
    if (ci >= 0) {
      // a step chosen by the testing knob fails as a HIP error would, before anything is enqueued
      rc = e->cfg.fault_step && ++S->planned == e->cfg.fault_step ? WF_E_DEVICE : serve_enqueue(e, S, ci, all, inits);
      if (rc == WF_OK) S->next_ctx ^= 1;
    } else {
      rc = serve_finish_oldest(e, S);
    }
    if (rc != WF_OK) serve_fail(e, S, rc);
  }
}

static void serve_free(Server* S) {
  (void)hipFree(S->d_acc);
  (void)hipFree(S->d_cks);
  (void)hipFree(S->d_sumq);
  (void)hipFree(S->d_nearv);
  for (ServeStep& c : S->step) {
    (void)hipFree(c.d_segs);
    (void)hipFree(c.d_runs);
    (void)hipFree(c.d_fills);
    (void)hipFree(c.d_ev);
    (void)hipFree(c.d_nev);
    (void)hipHostFree(c.h_acc);
    (void)hipHostFree(c.h_cks);
    (void)hipHostFree(c.h_sumq);
    (void)hipHostFree(c.h_nev);
    (void)hipHostFree(c.h_flags);
    if (c.done) (void)hipEventDestroy(c.done);
    for (auto& pass : c.tev)
      for (hipEvent_t& x : pass)
        if (x) (void)hipEventDestroy(x);
  }
  delete S;
}

extern "C" int wf_serve_start(wf_engine* e, u32 slots, u64 hbm_budget) {
  if (!e || slots == 0 || slots > 65536 || e->srv) return WF_E_ARG;
  HIPCHK(hipSetDevice(e->dev));
  Server* S = new Server();
  S->slots = slots;
  S->slot_ticket.assign(slots, 0);
  // bytes of pass buffers per number: prime regions (survivor lists in prove mode, at most
  // a tenth of the numbers for bucket_min >= 512) + buckets (large primes up to 2^32); two
  // buffer sets are in use at once
  const double SB = e->cfg.seg_bytes;
  const double pcap = std::max((double)prime_bound(30.0 * SB), e->cfg.prove == WF_PROVE_OFF ? 0.0 : 3.0 * SB);
  const double bcap = 8.0 * SB * 0.70 + 10.0 * std::sqrt(8.0 * SB * 0.70) + 1024.0;
  const double per_number = 2.0 * (pcap + bcap) * 4.0 / (30.0 * SB);
  if (!hbm_budget) hbm_budget = 16ull << 30;
  S->step_numbers = (u64)std::min(2.0e10, (double)hbm_budget / per_number);
  S->min_piece = std::max<u64>(30ull * e->cfg.seg_bytes, std::min<u64>(S->step_numbers / 64, 200000000ull));
  if (S->step_numbers < S->min_piece) { delete S; return WF_E_ARG; }
  S->nb = (u32)e->nb;
  const size_t nck = (size_t)slots * S->nb * sizeof(unsigned long long);
  bool ok = hipMalloc(&S->d_acc, slots * sizeof(DevAcc)) == hipSuccess && hipMalloc(&S->d_cks, nck) == hipSuccess &&
            hipMalloc(&S->d_sumq, nck) == hipSuccess && hipMalloc(&S->d_nearv, slots * sizeof(i64)) == hipSuccess;
  for (ServeStep& c : S->step)
    ok = ok && hipMalloc(&c.d_ev, SERVE_EVCAP * sizeof(DevEvent)) == hipSuccess &&
         hipMalloc(&c.d_nev, sizeof(unsigned long long)) == hipSuccess &&
         hipHostMalloc(&c.h_acc, slots * sizeof(DevAcc), 0) == hipSuccess &&
         hipHostMalloc(&c.h_cks, nck, 0) == hipSuccess && hipHostMalloc(&c.h_sumq, nck, 0) == hipSuccess &&
         hipHostMalloc(&c.h_nev, sizeof(unsigned long long), 0) == hipSuccess &&
         hipHostMalloc(&c.h_flags, sizeof(u32), 0) == hipSuccess &&
         hipEventCreateWithFlags(&c.done, hipEventDisableTiming) == hipSuccess;
  for (ServeStep& c : S->step)
    for (auto& pass : c.tev)
      for (hipEvent_t& x : pass) ok = ok && hipEventCreate(&x) == hipSuccess;
  if (!ok) {
    serve_free(S);
    return WF_E_NOMEM;
  }
  // set up once, before any step: every slot starts cleared
  if (hipMemset(S->d_acc, 0, slots * sizeof(DevAcc)) != hipSuccess || hipMemset(S->d_cks, 0, nck) != hipSuccess ||
      hipMemset(S->d_sumq, 0, nck) != hipSuccess || hipMemset(S->d_nearv, 0, slots * sizeof(i64)) != hipSuccess ||
      hipMemset(e->d_flags, 0, sizeof(u32)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    serve_free(S);
    return WF_E_DEVICE;
  }
  e->srv = S;
  S->th = std::thread(serve_loop, e);
  return WF_OK;
}

extern "C" int wf_serve_submit(wf_engine* e, u64 lo, u64 hi, int64_t near_T, u64* ticket) {
  if (!e || !e->srv || !ticket || (hi != 0 && hi <= lo)) return WF_E_ARG;
  Server* S = e->srv;
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->err) return S->err;
  // admission: a free slot, and the step budget still gives every request min_piece
  if (S->active >= S->slots || (u64)(S->active + 1) * S->min_piece > S->step_numbers) return WF_E_BUSY;
  u32 k = 0;
  while (S->slot_ticket[k]) k++;
  const u64 t = S->next_ticket++;
  ServeReq& q = S->reqs[t];
  q.lo = lo;
  q.hi = hi;
  q.next = lo;
  q.near_T = near_T;
  q.slot = k;
  q.scheduled = false;
  q.done = false;
  q.rc = WF_OK;
  memset(&q.res, 0, sizeof q.res);
  S->slot_ticket[k] = t;
  S->fresh.push_back({k, near_T});
  S->active++;
  *ticket = t;
  S->cv.notify_all();
  return WF_OK;
}

extern "C" int wf_serve_poll(wf_engine* e, u64 ticket, wf_result* out, wf_event* ev, u64 ev_cap, int* done) {
  if (!e || !e->srv || !out || !done || (ev_cap && !ev)) return WF_E_ARG;
  Server* S = e->srv;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return WF_E_ARG;
  ServeReq& q = it->second;
  *done = q.done ? 1 : 0;
  if (!q.done) return S->err;
  *out = q.res;
  for (u64 k = 0; k < q.ev.size() && k < ev_cap; k++) to_event(q.ev[k], 0, ev[k]);
  const int rc = q.rc;
  S->reqs.erase(it);
  return rc;
}

extern "C" int wf_serve_stop(wf_engine* e) {
  if (!e || !e->srv) return WF_E_ARG;
  Server* S = e->srv;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->th.joinable()) S->th.join();
  const int err = S->err;
  e->last = S->tim;
  e->last.total_ms = S->tim.sieve_ms + S->tim.test_ms;
  serve_free(S);
  e->srv = nullptr;
  return err;
}
