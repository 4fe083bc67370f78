// Bit-packed formats on the GPU: Polynomial::ToProto / CreateFromProto
// (crypto/polynomial.cc:562-590, 208-252) and the server image's
// BitPackCoefficients / BitUnpackCoefficients / BitUnpackSigned
// (crypto/pir/server/pir_server.cc:309-429). All of them lay value i of a
// stream on bits [i*bits, (i+1)*bits), least significant bit first within
// each byte; the signed unpack sign-extends from bit bits-1.
//
// On the device a stream is a run of 64-bit words (little-endian bytes, so
// bit j of the stream is bit j%64 of word j/64). Several equal streams (one
// per polynomial) sit at a fixed word pitch. Unpacking reads the one or two
// words a value straddles; packing gives each output word to one thread,
// which ORs in the (at most 64) values that overlap it. Both are bound by
// HBM bandwidth: 8 bytes per value on one side, bits/8 on the other.
#include <cstring>
#include <vector>

#include "pirgpu.h"
#include "server_internal.h"

namespace pirg {
namespace {

__device__ __forceinline__ u64 FieldMask(int bits) {
  return bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
}

// out[p][i] = value i of stream p (streams at `pitch` words, stream p's
// values from bit bit0 on). out_bytes: 8 (uint64), 4 (int32), 2 (int16).
// A value reads a second word only when its bits reach into it, so no word
// past the stream's last bit is touched.
__global__ void BitUnpackKernel(const u64* __restrict__ words, long long pitch, long long bit0,
                                int nstreams, long long len, int bits, int sign, int out_bytes,
                                void* __restrict__ out) {
  const long long total = static_cast<long long>(nstreams) * len;
  const u64 mask = FieldMask(bits);
  for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
       g += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long p = g / len;
    const long long i = g - p * len;
    const unsigned long long pos = static_cast<unsigned long long>(bit0) +
                                   static_cast<unsigned long long>(i) * bits;
    const u64* w = words + p * pitch + static_cast<long long>(pos >> 6);
    const int s = static_cast<int>(pos & 63);
    u64 v = w[0] >> s;
    if (s + bits > 64) v |= w[1] << (64 - s);
    v &= mask;
    if (sign && bits < 64 && ((v >> (bits - 1)) & 1ULL)) v |= ~mask;
    if (out_bytes == 8) {
      static_cast<u64*>(out)[g] = v;
    } else if (out_bytes == 4) {
      static_cast<int32_t*>(out)[g] = static_cast<int32_t>(v);
    } else {
      static_cast<int16_t*>(out)[g] = static_cast<int16_t>(v);
    }
  }
}

// words[p][w] = bits [64w, 64w + 64) of stream p, the low `bits` bits of
// vals[p][0..len) packed from bit 0; words past the stream's end are zero.
__global__ void BitPackKernel(const u64* __restrict__ vals, int nstreams, long long len,
                              int bits, u64* __restrict__ words, long long pitch) {
  const long long total = static_cast<long long>(nstreams) * pitch;
  const u64 mask = FieldMask(bits);
  for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
       g += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long p = g / pitch;
    const long long w = g - p * pitch;
    const long long lo = w * 64;
    const long long i0 = lo / bits;
    long long i1 = (lo + 63) / bits;
    if (i1 > len - 1) i1 = len - 1;
    const u64* src = vals + p * len;
    u64 acc = 0;
    for (long long i = i0; i <= i1; ++i) {
      const u64 v = src[i] & mask;
      const long long sh = i * bits - lo;  // in (-bits, 64)
      acc |= sh >= 0 ? (v << sh) : (v >> (-sh));
    }
    words[g] = acc;
  }
}

int CodecBlocks(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return static_cast<int>(b < 1 ? 1 : b);
}

}  // namespace

hipError_t LaunchBitUnpack(const u64* words, long long pitch, long long bit0, int nstreams,
                           long long len, int bits, bool sign, int out_bytes, void* out,
                           hipStream_t s) {
  if (nstreams <= 0 || len <= 0) return hipSuccess;
  if (bits < 1 || bits > 64 || (out_bytes != 8 && out_bytes != 4 && out_bytes != 2)) {
    return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(BitUnpackKernel, dim3(CodecBlocks(static_cast<long long>(nstreams) * len)),
                     dim3(256), 0, s, words, pitch, bit0, nstreams, len, bits, sign ? 1 : 0,
                     out_bytes, out);
  return hipGetLastError();
}

hipError_t LaunchBitPack(const u64* vals, int nstreams, long long len, int bits, u64* words,
                         long long pitch, hipStream_t s) {
  if (nstreams <= 0 || len <= 0) return hipSuccess;
  if (bits < 1 || bits > 64 || pitch * 64 < len * bits) return hipErrorInvalidValue;
  hipLaunchKernelGGL(BitPackKernel, dim3(CodecBlocks(static_cast<long long>(nstreams) * pitch)),
                     dim3(256), 0, s, vals, nstreams, len, bits, words, pitch);
  return hipGetLastError();
}

long long StreamWords(long long len, int bits) { return (len * bits + 63) / 64; }

namespace {

// CreateFromProto's checks (polynomial.cc:211-226), in its order.
Error CheckFromProto(int len, int log_modulus, size_t poly_bytes) {
  if (len <= 0 || (len & (len - 1)) != 0) return Invalid("Invalid length.");
  if (log_modulus <= 0 || log_modulus > 64) return Invalid("Invalid log_modulus.");
  if (poly_bytes == 0) return Invalid("Empty serialized polynomial.");
  const size_t expected = (static_cast<size_t>(len) * log_modulus + 7) / 8;
  if (poly_bytes != expected) return Invalid("Unexpected serialized size.");
  return Ok();
}

Error CheckToProto(int npolys, int len, int log_modulus) {
  if (log_modulus <= 0 || log_modulus > 64) return Invalid("Invalid log_modulus.");
  if (npolys < 0 || len <= 0) return Invalid("npolys must be >= 0 and len positive.");
  return Ok();
}

}  // namespace
}  // namespace pirg

using namespace pirg;

extern "C" {

int pirg_poly_to_proto(const uint64_t* coeffs, int npolys, int len, int log_modulus,
                       uint8_t* out, char* err, size_t errlen) {
  Error er = CheckToProto(npolys, len, log_modulus);
  if (!er.ok()) return Report(er, err, errlen);
  if (npolys == 0) return 0;
  er = CheckDevice();
  if (!er.ok()) return Report(er, err, errlen);
  const long long pitch = StreamWords(len, log_modulus);
  const size_t poly_bytes = (static_cast<size_t>(len) * log_modulus + 7) / 8;
  DevBuf vals, words;
  er = vals.alloc(static_cast<size_t>(npolys) * len * 8);
  if (er.ok()) er = words.alloc(static_cast<size_t>(npolys) * pitch * 8);
  if (!er.ok()) return Report(er, err, errlen);
  auto run = [&]() -> Error {
    PIRG_HIP(hipMemcpy(vals.p, coeffs, static_cast<size_t>(npolys) * len * 8,
                       hipMemcpyHostToDevice));
    PIRG_HIP(LaunchBitPack(vals.as<u64>(), npolys, len, log_modulus, words.as<u64>(), pitch,
                           nullptr));
    PIRG_HIP(hipMemcpy2D(out, poly_bytes, words.p, pitch * 8, poly_bytes, npolys,
                         hipMemcpyDeviceToHost));
    return Ok();
  };
  return Report(run(), err, errlen);
}

int pirg_poly_from_proto(const uint8_t* bytes, size_t poly_bytes, int npolys, int len,
                         int log_modulus, uint64_t* out, char* err, size_t errlen) {
  Error er = CheckFromProto(len, log_modulus, poly_bytes);
  if (!er.ok()) return Report(er, err, errlen);
  if (npolys < 0) return Report(Invalid("npolys must be >= 0."), err, errlen);
  if (npolys == 0) return 0;
  er = CheckDevice();
  if (!er.ok()) return Report(er, err, errlen);
  const long long pitch = StreamWords(len, log_modulus);
  DevBuf words, vals;
  er = words.alloc(static_cast<size_t>(npolys) * pitch * 8);
  if (er.ok()) er = vals.alloc(static_cast<size_t>(npolys) * len * 8);
  if (!er.ok()) return Report(er, err, errlen);
  auto run = [&]() -> Error {
    PIRG_HIP(hipMemset(words.p, 0, static_cast<size_t>(npolys) * pitch * 8));
    PIRG_HIP(hipMemcpy2D(words.p, pitch * 8, bytes, poly_bytes, poly_bytes, npolys,
                         hipMemcpyHostToDevice));
    PIRG_HIP(LaunchBitUnpack(words.as<u64>(), pitch, 0, npolys, len, log_modulus, false, 8,
                             vals.p, nullptr));
    PIRG_HIP(hipMemcpy(out, vals.p, static_cast<size_t>(npolys) * len * 8,
                       hipMemcpyDeviceToHost));
    return Ok();
  };
  return Report(run(), err, errlen);
}

int pirg_poly_to_proto_device(const uint64_t* d_coeffs, int npolys, int len, int log_modulus,
                              uint64_t* d_words, long long pitch_words, void* stream,
                              char* err, size_t errlen) {
  Error er = CheckToProto(npolys, len, log_modulus);
  if (!er.ok()) return Report(er, err, errlen);
  if (pitch_words < StreamWords(len, log_modulus)) {
    return Report(Invalid("pitch_words is below the stream length."), err, errlen);
  }
  auto run = [&]() -> Error {
    PIRG_HIP(LaunchBitPack(d_coeffs, npolys, len, log_modulus, d_words, pitch_words,
                           static_cast<hipStream_t>(stream)));
    return Ok();
  };
  return Report(run(), err, errlen);
}

int pirg_poly_from_proto_device(const uint64_t* d_words, long long pitch_words, int npolys,
                                int len, int log_modulus, uint64_t* d_coeffs, void* stream,
                                char* err, size_t errlen) {
  Error er = CheckFromProto(len, log_modulus, (static_cast<size_t>(len) * log_modulus + 7) / 8);
  if (!er.ok()) return Report(er, err, errlen);
  if (npolys < 0 || pitch_words < StreamWords(len, log_modulus)) {
    return Report(Invalid("pitch_words is below the stream length."), err, errlen);
  }
  auto run = [&]() -> Error {
    PIRG_HIP(LaunchBitUnpack(d_words, pitch_words, 0, npolys, len, log_modulus, false, 8,
                             d_coeffs, static_cast<hipStream_t>(stream)));
    return Ok();
  };
  return Report(run(), err, errlen);
}

}  // extern "C"
