// pirg_inner_product_fft_noalloc: Polynomial::InnerProductFft
// (crypto/polynomial_fft.cc:68-197) on device memory, for callers that must
// not allocate device memory (work under stream capture, serving loops that
// forbid hipMalloc). It runs the kernels of pirg_inner_product_fft
// (SpectraKernel, IpKernel, SumPairsKernel) and keeps everything beyond the
// caller's inputs and output in `scratch`, one static device array per GPU:
//
//   [0, kTablesBytes)        stage-ordered twiddle tables, forward then
//                            backward, for every N = 2^L, L = 2..13, written
//                            once per device at set-up and never again;
//   [kTablesBytes, +256)     IpKernel's pointer table (u spectra, v spectra,
//                            pair partials), written by a one-thread kernel
//                            on every call;
//   [kWorkOffset, end)       the call's u spectra, v spectra and pair
//                            partials, each at a 256-byte boundary.
//
// Every value a call reads from the work region is written earlier in the
// same call (the spectra and partials are stored, never accumulated), so no
// region needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "server_internal.h"

namespace pirg {
namespace {

// Largest working set of one call: the twiddle tables (512 KiB) plus
// n * (u chunks + v chunks) * 2d * 16 + pairs * d * 8 bytes. At d = 4096 a
// transform takes 128 KiB, so the 31.5 MiB work region holds up to 250 of
// them: ExternalProduct's six terms of one u and three v chunks need 24.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMinLogN = 2;   // d = 2
constexpr int kMaxLogN = 13;  // d = 4096
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t TableEntries(int L) { return (static_cast<size_t>(1) << L) - 1; }

// Byte offset of the forward (dir 0) or backward (dir 1) table of 2^L points.
size_t TwiddleOffset(int L, int dir) {
  size_t off = 0;
  for (int l = kMinLogN; l < L; ++l) off += 2 * TableEntries(l) * sizeof(double2);
  return off + static_cast<size_t>(dir) * TableEntries(L) * sizeof(double2);
}

size_t Align256(size_t x) { return (x + 255) / 256 * 256; }

const size_t kTablesBytes = Align256(TwiddleOffset(kMaxLogN + 1, 0));
const size_t kWorkOffset = kTablesBytes + 256;

// Per device: the address of its copy of `scratch` (from
// hipGetSymbolAddress, with that device current) once the tables are in it.
struct DeviceScratch {
  unsigned char* base = nullptr;
};
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

__global__ void PointerTableKernel(void** tab, void* us, void* vs, void* part) {
  tab[0] = us;
  tab[1] = vs;
  tab[2] = part;
}

// Once per device: finds the device's copy of `scratch`, uploads the twiddle
// tables through `s` and sets the FFT kernels' launch attributes.
Error SetUp(int dev, hipStream_t s, unsigned char** base) {
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceScratch& D = g_dev[dev];
  if (D.base != nullptr) {
    *base = D.base;
    return Ok();
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  PIRG_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) {
    return Error{PIRG_FAILED_PRECONDITION,
                 "the first call of pirg_inner_product_fft_noalloc on a device sets up "
                 "its tables and cannot run under stream capture"};
  }
  void* p = nullptr;
  PIRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  unsigned char* b = static_cast<unsigned char*>(p);
  for (int L = kMinLogN; L <= kMaxLogN; ++L) {
    std::vector<double2> fwd, bwd;
    Twiddles(1 << L, &fwd, &bwd);
    PIRG_HIP(hipMemcpyAsync(b + TwiddleOffset(L, 0), fwd.data(), fwd.size() * sizeof(double2),
                            hipMemcpyHostToDevice, s));
    PIRG_HIP(hipMemcpyAsync(b + TwiddleOffset(L, 1), bwd.data(), bwd.size() * sizeof(double2),
                            hipMemcpyHostToDevice, s));
    PIRG_HIP(hipStreamSynchronize(s));
    PIRG_HIP(ConfigureFftKernels(1 << L));
  }
  D.base = b;
  *base = b;
  return Ok();
}

Error Run(const uint64_t* u, const uint64_t* v, int n, int d, int u_bits, int v_bits,
          int chunk_bits, int u_is_signed, uint64_t* out, hipStream_t s) {
  if (n <= 0) return Invalid("Polynomial vectors cannot be empty.");
  if (d < 2 || d > 4096 || (d & (d - 1)) != 0) {
    return Invalid("Polynomial lengths or context mismatch.");
  }
  ChunkRule ur{}, vr{};
  Error er = MakeRules(n, d, u_bits, v_bits, chunk_bits, u_is_signed != 0, &ur, &vr);
  if (!er.ok()) return er;
  if (u == nullptr || v == nullptr || out == nullptr) {
    return Invalid("u, v and out must be device pointers.");
  }
  const int N = 2 * d;
  int log_d = 0;
  while ((1 << log_d) < d) ++log_d;
  const PairList pairs = MakePairs(ur, vr);
  const size_t us_bytes = Align256(static_cast<size_t>(n) * ur.num_chunks * N * sizeof(double2));
  const size_t vs_bytes = Align256(static_cast<size_t>(n) * vr.num_chunks * N * sizeof(double2));
  const size_t part_bytes = Align256(static_cast<size_t>(pairs.count) * d * 8);
  const size_t need = us_bytes + vs_bytes + part_bytes;
  if (need > kScratchBytes - kWorkOffset) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "InnerProductFft of %d terms at d = %d needs %zu bytes of working memory; "
                  "the no-allocation form has %zu",
                  n, d, need, kScratchBytes - kWorkOffset);
    return Invalid(msg);
  }
  er = CheckDevice();
  if (!er.ok()) return er;
  int dev = 0;
  PIRG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return Internal("device index out of range");
  unsigned char* base = nullptr;
  er = SetUp(dev, s, &base);
  if (!er.ok()) return er;

  const FftDims F{d, log_d, N, log_d + 1};
  const double2* twf = reinterpret_cast<const double2*>(base + TwiddleOffset(log_d + 1, 0));
  const double2* twb = reinterpret_cast<const double2*>(base + TwiddleOffset(log_d + 1, 1));
  void** tab = reinterpret_cast<void**>(base + kTablesBytes);
  double2* us = reinterpret_cast<double2*>(base + kWorkOffset);
  double2* vs = reinterpret_cast<double2*>(base + kWorkOffset + us_bytes);
  u64* part = reinterpret_cast<u64*>(base + kWorkOffset + us_bytes + vs_bytes);

  hipLaunchKernelGGL(PointerTableKernel, dim3(1), dim3(1), 0, s, tab, us, vs, part);
  PIRG_HIP(hipGetLastError());
  PIRG_HIP(LaunchSpectra(u, n, 1, d, ur, F, twf, us, s));
  PIRG_HIP(LaunchSpectra(v, n, 1, d, vr, F, twf, vs, s));
  PIRG_HIP(LaunchIp(reinterpret_cast<const double2* const*>(tab),
                    reinterpret_cast<const double2* const*>(tab + 1),
                    reinterpret_cast<u64* const*>(tab + 2), 1, 1, n, ur.num_chunks,
                    vr.num_chunks, pairs, F, twb, nullptr, s));
  PIRG_HIP(LaunchSumPairs(part, pairs.count, d, out, s));
  return Ok();
}

}  // namespace
}  // namespace pirg

extern "C" {

int pirg_inner_product_fft_noalloc(const uint64_t* u, const uint64_t* v, int n, int d,
                                   int u_bits, int v_bits, int chunk_bits, int u_is_signed,
                                   uint64_t* out, void* stream, char* err, size_t errlen) {
  const pirg::Error er = pirg::Run(u, v, n, d, u_bits, v_bits, chunk_bits, u_is_signed, out,
                                   static_cast<hipStream_t>(stream));
  return er.ok() ? 0 : pirg::Report(er, err, errlen);
}

}  // extern "C"
