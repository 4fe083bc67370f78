// Bit-packed formats on the GPU: Polynomial::ToProto / CreateFromProto
// (crypto/polynomial.cc:562-590, 208-252) and the server image's
// BitPackCoefficients / BitUnpackCoefficients / BitUnpackSigned
// (crypto/pir/server/pir_server.cc:309-429). All of them lay value i of a
// stream on bits [i*bits, (i+1)*bits), least significant bit first within
// each byte; the signed unpack sign-extends from bit bits-1.
//
// On the device a stream is a run of 64-bit words (little-endian bytes, so
// bit j of the stream is bit j%64 of word j/64). Several equal streams (one
// per polynomial) sit at a fixed word pitch. A workgroup owns a tile of 2048
// values: it stages the tile's words (unpacking) or values (packing) in LDS
// with coalesced loads, then each value reads the one or two words it
// straddles, or each output word ORs in the values that overlap it. Both are
// bound by HBM bandwidth: 8 bytes per value on one side, bits/8 on the other.
#include <cstring>
#include <vector>

#include "pirgpu.h"
#include "server_internal.h"

namespace pirg {
namespace {

__device__ __forceinline__ u64 FieldMask(int bits) {
  return bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
}

// Tiles of kTile values per workgroup (kTile / 256 per thread). A stream's
// tile t covers values [t*kTile, (t+1)*kTile) and, when the stream starts at
// bit 0, exactly the words [t*kTile*bits/64, (t+1)*kTile*bits/64).
constexpr int kTile = 2048;
constexpr int kTileThreads = 256;

// out[p][i] = value i of stream p (streams at `pitch` words, stream p's
// values from bit bit0 on). out_bytes: 8 (uint64), 4 (int32), 2 (int16).
// The tile's words are staged in LDS with coalesced loads; each value then
// reads its one or two words there. A value reads a second word only when
// its bits reach into it, so no word past the stream's last bit is touched.
// Grid: (tiles of a stream, streams).
__global__ __launch_bounds__(kTileThreads) void BitUnpackKernel(
    const u64* __restrict__ words, long long pitch, long long bit0, long long len, int bits,
    int sign, int out_bytes, void* __restrict__ out) {
  __shared__ u64 w[kTile + 2];
  const long long p = blockIdx.y;
  const long long v0 = static_cast<long long>(blockIdx.x) * kTile;
  const int cnt = static_cast<int>(len - v0 < kTile ? len - v0 : kTile);
  const unsigned long long first_bit = static_cast<unsigned long long>(bit0) +
                                       static_cast<unsigned long long>(v0) * bits;
  const unsigned long long last_bit = first_bit + static_cast<unsigned long long>(cnt) * bits - 1;
  const long long wa = static_cast<long long>(first_bit >> 6);
  const int nw = static_cast<int>((last_bit >> 6) - (first_bit >> 6)) + 1;
  const u64* src = words + p * pitch + wa;
  for (int i = threadIdx.x; i < nw; i += kTileThreads) w[i] = src[i];
  __syncthreads();
  const u64 mask = FieldMask(bits);
  const int base = static_cast<int>(first_bit & 63);
  const long long o = p * len + v0;
  for (int v = threadIdx.x; v < cnt; v += kTileThreads) {
    const unsigned pos = static_cast<unsigned>(base) + static_cast<unsigned>(v) * bits;
    const int s = pos & 63;
    const int j = pos >> 6;
    u64 x = w[j] >> s;
    if (s + bits > 64) x |= w[j + 1] << (64 - s);
    x &= mask;
    if (sign && bits < 64 && ((x >> (bits - 1)) & 1ULL)) x |= ~mask;
    if (out_bytes == 8) {
      static_cast<u64*>(out)[o + v] = x;
    } else if (out_bytes == 4) {
      static_cast<int32_t*>(out)[o + v] = static_cast<int32_t>(x);
    } else {
      static_cast<int16_t*>(out)[o + v] = static_cast<int16_t>(x);
    }
  }
}

// words[p][w] = bits [64w, 64w + 64) of stream p: the low `bits` bits of
// vals[p][0..len) packed from bit 0; words from the stream's end up to the
// pitch are written as zero. The tile's values are staged in LDS with
// coalesced loads; each thread then ORs together the values overlapping
// each of its output words. Grid: (tiles of a stream, streams).
__global__ __launch_bounds__(kTileThreads) void BitPackKernel(
    const u64* __restrict__ vals, long long len, int bits, u64* __restrict__ words,
    long long pitch) {
  __shared__ u64 v[kTile];
  const long long p = blockIdx.y;
  const long long v0 = static_cast<long long>(blockIdx.x) * kTile;
  const int cnt = static_cast<int>(len - v0 < kTile ? len - v0 : kTile);
  const u64* src = vals + p * len + v0;
  const u64 mask = FieldMask(bits);
  for (int i = threadIdx.x; i < cnt; i += kTileThreads) v[i] = src[i] & mask;
  __syncthreads();
  // This tile's words: [v0*bits/64, ceil((v0 + cnt)*bits/64)); kTile*bits is
  // a multiple of 64, so tiles own disjoint whole words.
  const long long w0 = v0 * bits / 64;
  const int nw = static_cast<int>((static_cast<long long>(cnt) * bits + 63) / 64);
  u64* dst = words + p * pitch + w0;
  for (int k = threadIdx.x; k < nw; k += kTileThreads) {
    const int lo = k * 64;
    const int i0 = lo / bits;
    int i1 = (lo + 63) / bits;
    if (i1 > cnt - 1) i1 = cnt - 1;
    u64 acc = 0;
    for (int i = i0; i <= i1; ++i) {
      const int sh = i * bits - lo;  // in (-bits, 64)
      acc |= sh >= 0 ? (v[i] << sh) : (v[i] >> (-sh));
    }
    dst[k] = acc;
  }
  // Words past the stream, up to the pitch: zero (written by the last tile).
  if (v0 + cnt == len) {
    const long long used = (len * bits + 63) / 64;
    for (long long k = used + threadIdx.x; k < pitch; k += kTileThreads) {
      words[p * pitch + k] = 0;
    }
  }
}

}  // namespace

hipError_t LaunchBitUnpack(const u64* words, long long pitch, long long bit0, int nstreams,
                           long long len, int bits, bool sign, int out_bytes, void* out,
                           hipStream_t s) {
  if (nstreams <= 0 || len <= 0) return hipSuccess;
  if (bits < 1 || bits > 64 || (out_bytes != 8 && out_bytes != 4 && out_bytes != 2) ||
      nstreams > 65535) {
    return hipErrorInvalidValue;
  }
  const long long tiles = (len + kTile - 1) / kTile;
  if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(BitUnpackKernel, dim3(static_cast<unsigned>(tiles), nstreams),
                     dim3(kTileThreads), 0, s, words, pitch, bit0, len, bits, sign ? 1 : 0,
                     out_bytes, out);
  return hipGetLastError();
}

hipError_t LaunchBitPack(const u64* vals, int nstreams, long long len, int bits, u64* words,
                         long long pitch, hipStream_t s) {
  if (nstreams <= 0 || len <= 0) return hipSuccess;
  if (bits < 1 || bits > 64 || pitch * 64 < len * bits || nstreams > 65535) {
    return hipErrorInvalidValue;
  }
  const long long tiles = (len + kTile - 1) / kTile;
  if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(BitPackKernel, dim3(static_cast<unsigned>(tiles), nstreams),
                     dim3(kTileThreads), 0, s, vals, len, bits, words, pitch);
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
    CallStream cs;
    PIRG_HIP(cs.create());
    PIRG_HIP(CopySync(vals.p, coeffs, static_cast<size_t>(npolys) * len * 8,
                      hipMemcpyHostToDevice, cs.s));
    PIRG_HIP(LaunchBitPack(vals.as<u64>(), npolys, len, log_modulus, words.as<u64>(), pitch,
                           cs.s));
    PIRG_HIP(hipMemcpy2DAsync(out, poly_bytes, words.p, pitch * 8, poly_bytes, npolys,
                              hipMemcpyDeviceToHost, cs.s));
    PIRG_HIP(hipStreamSynchronize(cs.s));
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
    CallStream cs;
    PIRG_HIP(cs.create());
    PIRG_HIP(hipMemsetAsync(words.p, 0, static_cast<size_t>(npolys) * pitch * 8, cs.s));
    PIRG_HIP(hipMemcpy2DAsync(words.p, pitch * 8, bytes, poly_bytes, poly_bytes, npolys,
                              hipMemcpyHostToDevice, cs.s));
    PIRG_HIP(LaunchBitUnpack(words.as<u64>(), pitch, 0, npolys, len, log_modulus, false, 8,
                             vals.p, cs.s));
    PIRG_HIP(CopySync(out, vals.p, static_cast<size_t>(npolys) * len * 8,
                      hipMemcpyDeviceToHost, cs.s));
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
