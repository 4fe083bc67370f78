// Calibration kernels for measurement only (libpirgpu_probes.so, declared in
// include/pirgpu_probes.h): a plain copy of known size that scales the HBM
// byte counters, a plain streaming read that gives the achievable read
// bandwidth, and a register-only int8 MFMA stream that gives the achievable
// MFMA rate. None of them is part of the answer path.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/pirgpu_probes.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

// Plain streaming copy (one 16-byte load and store per element).
__global__ void CopyProbeKernel(const int4* __restrict__ src, int4* __restrict__ dst,
                                long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    dst[i] = src[i];
  }
}

// Plain streaming read: `passes` sweeps over the buffer, every 16-byte
// element loaded once per sweep (8 non-temporal loads in flight per thread)
// and folded by XOR; one word per thread is written.
__global__ __launch_bounds__(256) void ReadProbeKernel(const v4i* __restrict__ src,
                                                       long long n, int passes,
                                                       int* __restrict__ out) {
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  int acc = 0;
  for (int pass = 0; pass < passes; ++pass) {
    long long i = tid;
    for (; i + 7 * stride < n; i += 8 * stride) {
      v4i v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    for (; i < n; i += stride) {
      const v4i v = __builtin_nontemporal_load(src + i);
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
  }
  out[tid] = acc;
}

// Register-only v_mfma_i32_16x16x64_i8 stream: 16 independent accumulators
// per wave, `iters` rounds; one output word per thread keeps the work
// observable.
__global__ __launch_bounds__(256) void MfmaProbeKernel(int iters, int* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  v4i acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = v4i{0, 0, 0, 0};
  const v4i a = v4i{lane, 1, 2, 3};
  const v4i b = v4i{3, lane, 1, 2};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[j], 0, 0, 0);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__device__ __forceinline__ unsigned ProbeHash(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// The same stream with random int8 operands: eight operand pairs per lane,
// hashed from the thread index and rotated across the accumulators, so that
// consecutive MFMAs see different random bytes as the scans' do. The matrix
// cores' clock drops with the operands' bit activity, so this rate, not the
// near-constant one above, is what a scan over random data can approach.
__global__ __launch_bounds__(256) void MfmaRandomProbeKernel(int iters, int* __restrict__ out) {
  constexpr int kSets = 8;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  v4i a[kSets], b[kSets];
#pragma unroll
  for (int s = 0; s < kSets; ++s) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      a[s][e] = static_cast<int>(ProbeHash(t * 64u + s * 8u + e));
      b[s][e] = static_cast<int>(ProbeHash(t * 64u + s * 8u + 4u + e) ^ 0x5bd1e995u);
    }
  }
  v4i acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = v4i{0, 0, 0, 0};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
      acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[j % kSets], b[(j / 2) % kSets], acc[j], 0,
                                                     0, 0);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[t] = s;
}

constexpr int kReadBlocks = 2048;

// Device buffers, a non-blocking stream and two events, released on return.
struct Ctx {
  void* a = nullptr;
  void* b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Ctx() {
    if (s != nullptr) (void)hipStreamSynchronize(s);
    if (a != nullptr) (void)hipFree(a);
    if (b != nullptr) (void)hipFree(b);
    if (e0 != nullptr) (void)hipEventDestroy(e0);
    if (e1 != nullptr) (void)hipEventDestroy(e1);
    if (s != nullptr) (void)hipStreamDestroy(s);
  }
  bool Init(size_t abytes, size_t bbytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return false;
    return hipMalloc(&a, std::max<size_t>(abytes, 16)) == hipSuccess &&
           hipMalloc(&b, std::max<size_t>(bbytes, 16)) == hipSuccess &&
           hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  }
  // Milliseconds between e0 and e1 once e1 has completed.
  float Elapsed() {
    float t = 0;
    if (hipEventSynchronize(e1) != hipSuccess) return -1.0f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return -1.0f;
    return t;
  }
};

int MfmaProbe(bool random, int blocks, int iters, double* ms, double* tops) {
  if (blocks <= 0 || iters <= 0) return 3;
  Ctx c;
  if (!c.Init(static_cast<size_t>(blocks) * 256 * 4, 16)) return 14;
  int* out = static_cast<int*>(c.a);
  auto kernel = random ? MfmaRandomProbeKernel : MfmaProbeKernel;
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, c.s, iters, out);
  (void)hipEventRecord(c.e0, c.s);
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, c.s, iters, out);
  (void)hipEventRecord(c.e1, c.s);
  const float t = c.Elapsed();
  if (t < 0 || hipGetLastError() != hipSuccess) return 13;
  if (ms) *ms = t;
  // 4 waves per block, 16 MFMAs per round, 16 x 16 x 64 x 2 ops each.
  const double ops = static_cast<double>(blocks) * 4 * iters * 16 * 32768.0;
  if (tops) *tops = ops / (t * 1e-3) / 1e12;
  return 0;
}

}  // namespace

extern "C" {

int pirg_copy_probe(size_t bytes, int repeats, double* ms) {
  if (ms == nullptr || repeats < 1) return 3;
  Ctx c;
  if (!c.Init(bytes, bytes)) return 14;
  if (hipMemsetAsync(c.a, 1, bytes, c.s) != hipSuccess) return 13;
  const long long n = static_cast<long long>(bytes / 16);
  auto copy = [&] {
    hipLaunchKernelGGL(CopyProbeKernel, dim3(4096), dim3(256), 0, c.s,
                       static_cast<const int4*>(c.a), static_cast<int4*>(c.b), n);
  };
  copy();
  (void)hipEventRecord(c.e0, c.s);
  for (int i = 0; i < repeats; ++i) copy();
  (void)hipEventRecord(c.e1, c.s);
  const float t = c.Elapsed();
  if (t < 0 || hipGetLastError() != hipSuccess) return 13;
  *ms = t / repeats;
  return 0;
}

int pirg_read_probe(size_t bytes, int repeats, double* ms) {
  if (ms == nullptr || repeats < 1) return 3;
  Ctx c;
  if (!c.Init(bytes, static_cast<size_t>(kReadBlocks) * 256 * 4)) return 14;
  if (hipMemsetAsync(c.a, 1, bytes, c.s) != hipSuccess) return 13;
  const long long n = static_cast<long long>(bytes / 16);
  auto read = [&](int passes) {
    hipLaunchKernelGGL(ReadProbeKernel, dim3(kReadBlocks), dim3(256), 0, c.s,
                       static_cast<const v4i*>(c.a), n, passes, static_cast<int*>(c.b));
  };
  read(1);
  (void)hipEventRecord(c.e0, c.s);
  read(repeats);
  (void)hipEventRecord(c.e1, c.s);
  const float t = c.Elapsed();
  if (t < 0 || hipGetLastError() != hipSuccess) return 13;
  *ms = t / repeats;
  return 0;
}


int pirg_mfma_probe(int blocks, int iters, double* ms, double* tops) {
  return MfmaProbe(false, blocks, iters, ms, tops);
}

int pirg_mfma_probe_random(int blocks, int iters, double* ms, double* tops) {
  return MfmaProbe(true, blocks, iters, ms, tops);
}

}  // extern "C"
