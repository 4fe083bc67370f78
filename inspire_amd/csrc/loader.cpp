// PirServer::LoadFromBuffer (crypto/pir/server/pir_server.cc:594-793): the
// server image written by SaveToFile (:433-592), read in the same order:
//   1. combined DB matrix: rows, physical cols (int32), condensed (uint8),
//      original cols (int32), element count (uint64), the raw elements
//      (condensed 64-bit words, or DB values when not condensed);
//   2. combined pack matrix: rows, cols, physical cols (int32), condensed
//      (uint8), original cols, bits per entry (int32), byte count (uint64)
//      and the entries bit-packed as signed values (BitPackSigned);
//   3. log modulus (int32), shard count (int32); per shard the slice count
//      (int32); per slice a_tilde_agg as (length, byte count, bits) and
//      t_vec_h as (count, then per polynomial length, byte count, bits);
//   4. second_dim_a count (uint64), then per polynomial (length, byte count,
//      bits).
// All integers little-endian; polynomials bit-packed at log-modulus bits
// (BitPackCoefficients). The walk here only records where each record lies;
// the pack matrix and the polynomials are unpacked by the GPU.
#include <cstring>
#include <vector>

#include "pirgpu.h"
#include "server_internal.h"

namespace pirg {
namespace {

// LoadFromBuffer's BufferReader: a read past the end leaves the offset where
// it is and marks the reader bad.
class ImageReader {
 public:
  ImageReader(const uint8_t* buf, size_t size) : buf_(buf), size_(size) {}
  template <typename T>
  T Get() {
    T v{};
    if (off_ + sizeof(T) > size_) {
      ok_ = false;
      return v;
    }
    std::memcpy(&v, buf_ + off_, sizeof(T));
    off_ += sizeof(T);
    return v;
  }
  // Skips n bytes and returns where they start (null when past the end).
  const uint8_t* Take(uint64_t n) {
    if (n > size_ - off_) {
      ok_ = false;
      return nullptr;
    }
    const uint8_t* p = buf_ + off_;
    off_ += n;
    return p;
  }
  bool ok() const { return ok_; }

 private:
  const uint8_t* buf_;
  size_t size_;
  size_t off_ = 0;
  bool ok_ = true;
};

Error ReadError() { return Error{13, "Error reading from buffer"}; }

struct PolyRecord {
  const uint8_t* bytes;
  uint64_t nbytes;
};

// A polynomial record (length, byte count, bytes) of d coefficients at
// logm bits.
Error ReadPoly(ImageReader& in, int d, int logm, std::vector<PolyRecord>* out) {
  const uint64_t len = in.Get<uint64_t>();
  const uint64_t nb = in.Get<uint64_t>();
  const uint8_t* p = in.Take(nb);
  if (!in.ok()) return ReadError();
  if (len != static_cast<uint64_t>(d)) {
    return Invalid("image polynomial has " + std::to_string(len) + " coefficients, expected " +
                   std::to_string(d) + ".");
  }
  // BitPackCoefficients copies whole words at 64 bits: the same byte count.
  if (nb != (static_cast<uint64_t>(d) * logm + 7) / 8) {
    return Invalid("image polynomial byte count does not match its length.");
  }
  out->push_back(PolyRecord{p, nb});
  return Ok();
}

Error LoadFromBuffer(const pirg_params* params, const uint8_t* buffer, size_t size,
                     std::unique_ptr<pirg_server>* out) {
  if (buffer == nullptr && size > 0) return Invalid("buffer is null.");
  pirg_server probe;
  Error er = DeriveServer(*params, &probe);
  if (!er.ok()) return er;
  const int d = probe.d, t = probe.t, ell = probe.ell, k = probe.k, e = probe.e, M = probe.M;
  ImageReader in(buffer, size);

  // 1. DB.
  const int32_t db_rows = in.Get<int32_t>();
  const int32_t db_phys = in.Get<int32_t>();
  const uint8_t db_condensed = in.Get<uint8_t>();
  const int32_t db_orig = in.Get<int32_t>();
  const uint64_t db_elems = in.Get<uint64_t>();
  if (!in.ok()) return ReadError();
  const int per = params->db_bytes == 2 ? 4 : 8;
  // Logical columns; the non-condensed elements are DB values whose bytes,
  // row-major, equal the condensed words when the columns fill whole words.
  const int db_cols = db_condensed != 0 ? db_orig : db_phys;
  const uint64_t elem_bytes = db_condensed != 0 ? 8 : static_cast<uint64_t>(params->db_bytes);
  if (db_rows != M) return Invalid("image DB matrix must have ell*t*d rows.");
  if (db_cols <= 0 || db_cols % per != 0 ||
      (db_condensed != 0 && db_phys != db_cols / per) ||
      db_elems != static_cast<uint64_t>(db_rows) * db_phys) {
    return Invalid("image DB matrix layout does not match its dimensions.");
  }
  const uint8_t* db_bytes = in.Take(db_elems * elem_bytes);
  if (!in.ok()) return ReadError();

  // 2. Pack matrix.
  const int32_t pk_rows = in.Get<int32_t>();
  const int32_t pk_cols = in.Get<int32_t>();
  (void)in.Get<int32_t>();  // physical columns
  (void)in.Get<uint8_t>();  // condensed: the entries are written in logical order either way
  (void)in.Get<int32_t>();  // original columns
  const int32_t pk_bits = in.Get<int32_t>();
  const uint64_t pk_len = in.Get<uint64_t>();
  const uint8_t* pk_bytes = in.Take(pk_len);
  if (!in.ok()) return ReadError();
  if (pk_rows != M || pk_cols != k * d) {
    return Invalid("image pack matrix must be ell*t*d x k*d.");
  }
  if (pk_bits < 1 || pk_bits > 64 ||
      pk_len < (static_cast<uint64_t>(pk_rows) * pk_cols * pk_bits + 7) / 8) {
    return Invalid("image pack matrix bit width or size is inconsistent.");
  }

  // 3. Preprocessed outputs.
  const int32_t logm = in.Get<int32_t>();
  const int32_t nshards = in.Get<int32_t>();
  if (!in.ok()) return ReadError();
  const int want_logm = params->log_q >= 64 ? 64 : params->log_q;
  if (logm != want_logm) return Invalid("image log modulus does not match the parameters.");
  if (nshards != ell) return Invalid("preprocessed_outputs size mismatch.");
  std::vector<PolyRecord> a_recs, t_recs, s_recs;
  for (int s = 0; s < ell; ++s) {
    const int32_t nslices = in.Get<int32_t>();
    if (!in.ok()) return ReadError();
    if (nslices != t) return Invalid("Preprocessed outputs must be of length t.");
    for (int i = 0; i < t; ++i) {
      er = ReadPoly(in, d, logm, &a_recs);
      if (!er.ok()) return er;
      const uint64_t tv = in.Get<uint64_t>();
      if (!in.ok()) return ReadError();
      if (tv != static_cast<uint64_t>(k)) return Invalid("t_vec_h must hold k polynomials.");
      for (int j = 0; j < k; ++j) {
        er = ReadPoly(in, d, logm, &t_recs);
        if (!er.ok()) return er;
      }
    }
  }
  // 4. second_dim_a.
  const uint64_t nsda = in.Get<uint64_t>();
  if (!in.ok()) return ReadError();
  if (nsda != static_cast<uint64_t>(t > 1 ? 2 * e : 0)) {
    return Invalid(
        "second_dim_query_a must have length 2 * PolyEvalGadgetParams().num_digits.");
  }
  for (uint64_t j = 0; j < nsda; ++j) {
    er = ReadPoly(in, d, logm, &s_recs);
    if (!er.ok()) return er;
  }

  // Polynomials -> device, in the order [a_tilde][t_vec_h][second_dim_a].
  er = CheckDevice();
  if (!er.ok()) return er;
  std::vector<const PolyRecord*> all;
  for (const auto& r : a_recs) all.push_back(&r);
  for (const auto& r : t_recs) all.push_back(&r);
  for (const auto& r : s_recs) all.push_back(&r);
  const long long pitch = StreamWords(d, logm);
  std::vector<u64> staged(all.size() * static_cast<size_t>(pitch), 0);
  for (size_t i = 0; i < all.size(); ++i) {
    std::memcpy(staged.data() + i * pitch, all[i]->bytes, all[i]->nbytes);
  }
  DevBuf words, polys;
  er = words.alloc(staged.size() * 8);
  if (!er.ok()) return er;
  er = polys.alloc(all.size() * static_cast<size_t>(d) * 8);
  if (!er.ok()) return er;
  CallStream cs;
  PIRG_HIP(cs.create());
  PIRG_HIP(CopySync(words.p, staged.data(), staged.size() * 8, hipMemcpyHostToDevice, cs.s));
  PIRG_HIP(LaunchBitUnpack(words.as<u64>(), pitch, 0, static_cast<int>(all.size()), d, logm,
                           false, 8, polys.p, cs.s));
  PIRG_HIP(hipStreamSynchronize(cs.s));
  const u64* dev = polys.as<u64>();
  const size_t lt = static_cast<size_t>(ell) * t;
  const u64* a_dev = dev;
  const u64* t_dev = dev + lt * d;
  const u64* s_dev = nsda > 0 ? dev + lt * (1 + k) * d : nullptr;

  // The DB words: the image's bytes are not 8-aligned, and the upload copies
  // them through a staging buffer anyway.
  const uint64_t* db_words = reinterpret_cast<const uint64_t*>(db_bytes);
  return CreateFromImage(params, db_words, db_cols, PackedPack{pk_bytes, pk_len, pk_bits},
                         a_dev, t_dev, s_dev, out);
}

}  // namespace
}  // namespace pirg

using namespace pirg;

extern "C" int pirg_server_load_from_buffer(const pirg_params* params, const uint8_t* buffer,
                                            size_t size, pirg_server** out, char* err,
                                            size_t errlen) {
  if (params == nullptr || out == nullptr) {
    return Report(Invalid("params and out must not be null."), err, errlen);
  }
  std::unique_ptr<pirg_server> S;
  Error er = LoadFromBuffer(params, buffer, size, &S);
  if (!er.ok()) return Report(er, err, errlen);
  *out = S.release();
  return 0;
}
