// wf_probe.hip -- small measurement helpers for the profiles of this engine.
//
//   wf_probe valu         VALU issue rates on gfx950 (lane-ops per clock per CU) of the
//                         instructions the kernels are built from
//   wf_probe copy BYTES   a plain copy kernel of BYTES bytes (calibrates the FETCH_SIZE /
//                         WRITE_SIZE counters of rocprofv3 --pmc against a known traffic)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CH 8
#define ITERS 2048

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      exit(1);                                                               \
    }                                                                        \
  } while (0)

// Each kernel runs CH independent chains whose next value depends on the previous one
// through the measured instruction; the xor with a per-iteration value keeps the compiler
// from folding the chain.
__global__ void k_mullo(const uint32_t* in, uint32_t* out) {
  uint32_t x[CH], y = in[threadIdx.x & 63] | 1;
  for (int c = 0; c < CH; c++) x[c] = in[(threadIdx.x + c) & 63];
  for (int i = 0; i < ITERS; i++)
#pragma unroll
    for (int c = 0; c < CH; c++) x[c] = (x[c] * y) ^ (uint32_t)i;
  uint32_t s = 0;
  for (int c = 0; c < CH; c++) s ^= x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
__global__ void k_mad64(const uint32_t* in, uint32_t* out) {
  uint64_t a[CH];
  const uint32_t y = in[threadIdx.x & 63] | 1;
  for (int c = 0; c < CH; c++) a[c] = in[(threadIdx.x + c) & 63];
  for (int i = 0; i < ITERS; i++)
#pragma unroll
    for (int c = 0; c < CH; c++) a[c] = (uint64_t)(uint32_t)a[c] * y + (a[c] >> 32);
  uint64_t s = 0;
  for (int c = 0; c < CH; c++) s ^= a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(s ^ (s >> 32));
}
__global__ void k_fma64(const uint32_t* in, uint32_t* out) {
  double x[CH];
  const double y = 1.0 + in[threadIdx.x & 63] * 1e-12, z = 1e-9;
  for (int c = 0; c < CH; c++) x[c] = in[(threadIdx.x + c) & 63];
  for (int i = 0; i < ITERS; i++)
#pragma unroll
    for (int c = 0; c < CH; c++) x[c] = __builtin_fma(x[c], y, z);
  double s = 0;
  for (int c = 0; c < CH; c++) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int64_t)s;
}
__global__ void k_add64(const uint32_t* in, uint32_t* out) {
  double x[CH];
  const double y = 1e-9;
  for (int c = 0; c < CH; c++) x[c] = in[(threadIdx.x + c) & 63];
  for (int i = 0; i < ITERS; i++)
#pragma unroll
    for (int c = 0; c < CH; c++) x[c] = x[c] + y;
  double s = 0;
  for (int c = 0; c < CH; c++) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int64_t)s;
}
__global__ void k_rndne64(const uint32_t* in, uint32_t* out) {
  double x[CH];
  for (int c = 0; c < CH; c++) x[c] = in[(threadIdx.x + c) & 63] + 0.25;
  for (int i = 0; i < ITERS; i++)
#pragma unroll
    for (int c = 0; c < CH; c++) x[c] = __builtin_rint(x[c]) + 0.25;
  double s = 0;
  for (int c = 0; c < CH; c++) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int64_t)s;
}

typedef void (*kfn)(const uint32_t*, uint32_t*);

static int valu() {
  hipDeviceProp_t pr;
  CHECK(hipGetDeviceProperties(&pr, 0));
  const int nblk = pr.multiProcessorCount * 16, nth = 256;
  uint32_t *in, *out;
  CHECK(hipMalloc(&in, 64 * 4));
  CHECK(hipMalloc(&out, (size_t)nblk * nth * 4));
  uint32_t h[64];
  for (int i = 0; i < 64; i++) h[i] = 0x9E3779B9u * (i + 1);
  CHECK(hipMemcpy(in, h, sizeof h, hipMemcpyHostToDevice));
  struct { const char* n; kfn f; double ops; } ks[] = {
      {"v_mul_lo_u32 (+v_xor)", k_mullo, 2}, {"v_mad_u64_u32", k_mad64, 1}, {"v_fma_f64", k_fma64, 1},
      {"v_add_f64", k_add64, 1}, {"v_rndne_f64 (+v_add_f64)", k_rndne64, 2}};
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  printf("device %s, %d CUs, %d MHz\n", pr.gcnArchName, pr.multiProcessorCount, pr.clockRate / 1000);
  for (auto& k : ks) {
    hipLaunchKernelGGL(k.f, dim3(nblk), dim3(nth), 0, 0, in, out);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k.f, dim3(nblk), dim3(nth), 0, 0, in, out);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms;
    CHECK(hipEventElapsedTime(&ms, a, b));
    const double ops = 5.0 * nblk * nth * (double)ITERS * CH * k.ops;
    printf("%-26s %.3e lane-ops/s  %6.2f lane-ops/clk/CU\n", k.n, ops / (ms * 1e-3),
           ops / (ms * 1e-3) / pr.multiProcessorCount / (pr.clockRate * 1e3));
  }
  return 0;
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

static int copy(size_t bytes) {
  const size_t n = bytes / sizeof(uint4);
  uint4 *s, *d;
  CHECK(hipMalloc(&s, n * sizeof(uint4)));
  CHECK(hipMalloc(&d, n * sizeof(uint4)));
  CHECK(hipMemset(s, 1, n * sizeof(uint4)));
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, s, d, n);
  CHECK(hipDeviceSynchronize());
  printf("copied %zu bytes (read %zu, write %zu)\n", n * sizeof(uint4), n * sizeof(uint4), n * sizeof(uint4));
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && !strcmp(argv[1], "valu")) return valu();
  if (argc >= 3 && !strcmp(argv[1], "copy")) return copy(strtoull(argv[2], 0, 10));
  fprintf(stderr, "usage: wf_probe valu | wf_probe copy BYTES\n");
  return 2;
}
