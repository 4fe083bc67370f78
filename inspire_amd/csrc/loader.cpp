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
//
// SaveToFile writes the same image from a resident server: the DB words and
// the pack entries are rebuilt from their int8-plane tiles on the GPU and
// bit-packed by the codec's encoder in row chunks, the polynomials are
// bit-packed from the raw copies the server keeps, and the chunks stream
// through a pinned buffer to the file (or a caller's buffer).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
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

// ---------------------------------------------------------------------------
// PirServer::SaveToFile (crypto/pir/server/pir_server.cc:433-592).

// Where the image goes: a file, a caller's buffer, or nowhere (counting).
class ImageSink {
 public:
  explicit ImageSink(FILE* f) : f_(f) {}
  ImageSink(uint8_t* buf, size_t cap) : buf_(buf), cap_(cap) {}
  bool Write(const void* p, size_t n) {
    if (n == 0) return true;
    if (f_ != nullptr) {
      if (std::fwrite(p, 1, n, f_) != n) return false;
    } else if (buf_ != nullptr) {
      if (n > cap_ - off_) return false;
      std::memcpy(buf_ + off_, p, n);
    }
    off_ += n;
    return true;
  }
  template <typename T>
  bool Val(T v) {
    return Write(&v, sizeof(T));
  }
  size_t written() const { return off_; }

 private:
  FILE* f_ = nullptr;
  uint8_t* buf_ = nullptr;
  size_t cap_ = 0;
  size_t off_ = 0;
};

struct ImageShape {
  int M, K, phys_db, pk_cols, pk_phys, pk_condensed, pk_bits, logm, nsda;
  uint64_t db_words, pk_bytes, poly_bytes;
};

// pir_server.cc:441-483: the header fields of the two matrices.
ImageShape ShapeOf(const pirg_server* S) {
  ImageShape sh{};
  sh.M = S->M;
  sh.K = S->K;
  sh.phys_db = S->P.db_bytes == 2 ? S->K / 4 : S->K / 8;
  sh.db_words = static_cast<uint64_t>(sh.M) * sh.phys_db;
  sh.pk_cols = S->k * S->d;
  // Matrix<int32_t> is condensed (two entries per word); Matrix<int16_t> and
  // Matrix<uint64_t> are plain (lwes_to_rlwe.cc:228-234).
  sh.pk_condensed = S->mat_src_bytes == 4 ? 1 : 0;
  sh.pk_phys = sh.pk_condensed ? (sh.pk_cols + 1) / 2 : sh.pk_cols;
  int log2_d = 0;
  while ((1 << log2_d) < S->d) ++log2_d;  // absl::bit_width(d - 1) for d >= 2
  sh.pk_bits = S->P.pack_log_digit + log2_d;
  sh.pk_bytes = (static_cast<uint64_t>(sh.M) * sh.pk_cols * sh.pk_bits + 7) / 8;
  sh.logm = S->log_q;  // RlweParams::LogModulus: 64 (or 32) for modulus 0
  sh.poly_bytes = (static_cast<uint64_t>(S->d) * sh.logm + 7) / 8;
  sh.nsda = S->t > 1 ? 2 * S->e : 0;
  return sh;
}

uint64_t ImageBytes(const pirg_server* S) {
  const ImageShape sh = ShapeOf(S);
  const uint64_t lt = static_cast<uint64_t>(S->ell) * S->t;
  const uint64_t poly_rec = 16 + sh.poly_bytes;
  return 4 + 4 + 1 + 4 + 8 + sh.db_words * 8 +            // DB
         4 * 5 + 1 + 8 + sh.pk_bytes +                      // pack matrix
         4 + 4 + 4ULL * S->ell +                            // log modulus, shards, slices
         lt * (poly_rec + 8 + static_cast<uint64_t>(S->k) * poly_rec) +
         8 + static_cast<uint64_t>(sh.nsda) * poly_rec;     // second_dim_a
}

Error WriteFailed(const std::string& where) {
  return Error{PIRG_INTERNAL, "Error writing to file: " + where};
}

// Rows of a tiled matrix, chunk by chunk: tiles -> values (GPU) -> bit-packed
// at `bits` (GPU) -> pinned host -> sink. `total_bytes` of the bit stream are
// written (a chunk of a multiple of 64 rows ends on a word boundary).
Error WriteTiledRows(const int8_t* tiles, const LimbLayout& L, bool signed_top, u64 mask,
                     int bits, uint64_t total_bytes, hipStream_t s, ImageSink& out,
                     const std::string& where) {
  const size_t row_vals = static_cast<size_t>(L.cols);
  int rows_per = static_cast<int>(std::max<size_t>(64, (size_t(512) << 20) / (row_vals * 8)));
  rows_per = std::max(64, rows_per / 64 * 64);
  rows_per = std::min(rows_per, (L.rows + 63) / 64 * 64);
  const long long max_words = StreamWords(static_cast<long long>(rows_per) * L.cols, bits);
  DevBuf vals, words;
  HostBuf host;
  Error er = vals.alloc(static_cast<size_t>(rows_per) * row_vals * 8);
  if (!er.ok()) return er;
  if ((er = words.alloc(static_cast<size_t>(max_words) * 8)).code) return er;
  if ((er = host.alloc(static_cast<size_t>(max_words) * 8)).code) return er;
  uint64_t left = total_bytes;
  for (int r0 = 0; r0 < L.rows; r0 += rows_per) {
    const int nr = std::min(rows_per, L.rows - r0);
    const long long len = static_cast<long long>(nr) * L.cols;
    const long long nw = StreamWords(len, bits);
    PIRG_HIP(LaunchTilesToValues(tiles, L, r0, nr, signed_top, mask, vals.as<u64>(), s));
    PIRG_HIP(LaunchBitPack(vals.as<u64>(), 1, len, bits, words.as<u64>(), nw, s));
    PIRG_HIP(hipMemcpyAsync(host.p, words.p, static_cast<size_t>(nw) * 8,
                            hipMemcpyDeviceToHost, s));
    PIRG_HIP(hipStreamSynchronize(s));
    const uint64_t n = std::min<uint64_t>(left, static_cast<uint64_t>(nw) * 8);
    if (!out.Write(host.p, n)) return WriteFailed(where);
    left -= n;
  }
  if (left != 0) return Internal("image matrix stream is short");
  return Ok();
}

Error SaveImage(pirg_server* S, ImageSink& out, const std::string& where) {
  const ImageShape sh = ShapeOf(S);
  CallStream cs;
  PIRG_HIP(cs.create());
  // 1. The combined DB matrix, condensed (pir_server.cc:441-464).
  bool ok = out.Val<int32_t>(sh.M) && out.Val<int32_t>(sh.phys_db) && out.Val<uint8_t>(1) &&
            out.Val<int32_t>(sh.K) && out.Val<uint64_t>(sh.db_words);
  if (!ok) return WriteFailed(where);
  Error er = WriteTiledRows(S->db_tiles.as<int8_t>(), S->dbL, false, ~0ULL,
                            8 * S->P.db_bytes, sh.db_words * 8, cs.s, out, where);
  if (!er.ok()) return er;
  // 2. The combined pack matrix, BitPackSigned at pack_log_digit + log2(d)
  // bits (:466-516): each entry cast to its unsigned type, then masked.
  ok = out.Val<int32_t>(sh.M) && out.Val<int32_t>(sh.pk_cols) && out.Val<int32_t>(sh.pk_phys) &&
       out.Val<uint8_t>(static_cast<uint8_t>(sh.pk_condensed)) &&
       out.Val<int32_t>(sh.pk_cols) && out.Val<int32_t>(sh.pk_bits) &&
       out.Val<uint64_t>(sh.pk_bytes);
  if (!ok) return WriteFailed(where);
  const u64 type_mask = S->mat_src_bytes == 2 ? 0xFFFFULL
                        : S->mat_src_bytes == 4 ? 0xFFFFFFFFULL : ~0ULL;
  er = WriteTiledRows(S->pk_tiles.as<int8_t>(), S->pkL, true, type_mask, sh.pk_bits,
                      sh.pk_bytes, cs.s, out, where);
  if (!er.ok()) return er;
  // 3-4. The polynomials, BitPackCoefficients at log-modulus bits
  // (:518-583): a_tilde_agg, t_vec_h, second_dim_a, one stream each.
  const int d = S->d, k = S->k, lt = S->ell * S->t;
  const long long pitch = StreamWords(d, sh.logm);
  const int npolys = lt + lt * k + sh.nsda;
  DevBuf words;
  HostBuf host;
  if ((er = words.alloc(static_cast<size_t>(npolys) * pitch * 8)).code) return er;
  if ((er = host.alloc(static_cast<size_t>(npolys) * pitch * 8)).code) return er;
  u64* w = words.as<u64>();
  PIRG_HIP(LaunchBitPack(S->a_tilde.as<u64>(), lt, d, sh.logm, w, pitch, cs.s));
  PIRG_HIP(LaunchBitPack(S->t_raw.as<u64>(), lt * k, d, sh.logm, w + lt * pitch, pitch, cs.s));
  if (sh.nsda > 0) {
    PIRG_HIP(LaunchBitPack(S->sda_raw.as<u64>(), sh.nsda, d, sh.logm,
                           w + static_cast<long long>(lt) * (1 + k) * pitch, pitch, cs.s));
  }
  PIRG_HIP(hipMemcpyAsync(host.p, words.p, static_cast<size_t>(npolys) * pitch * 8,
                          hipMemcpyDeviceToHost, cs.s));
  PIRG_HIP(hipStreamSynchronize(cs.s));
  const uint8_t* hb = host.as<uint8_t>();
  auto poly = [&](int idx) {
    return out.Val<uint64_t>(d) && out.Val<uint64_t>(sh.poly_bytes) &&
           out.Write(hb + static_cast<size_t>(idx) * pitch * 8, sh.poly_bytes);
  };
  ok = out.Val<int32_t>(sh.logm) && out.Val<int32_t>(S->ell);
  for (int s = 0; ok && s < S->ell; ++s) {
    ok = out.Val<int32_t>(S->t);
    for (int i = 0; ok && i < S->t; ++i) {
      const int oi = s * S->t + i;
      ok = poly(oi) && out.Val<uint64_t>(k);
      for (int j = 0; ok && j < k; ++j) ok = poly(lt + oi * k + j);
    }
  }
  ok = ok && out.Val<uint64_t>(sh.nsda);
  for (int j = 0; ok && j < sh.nsda; ++j) ok = poly(lt + lt * k + j);
  if (!ok) return WriteFailed(where);
  return Ok();
}

}  // namespace
}  // namespace pirg

using namespace pirg;

extern "C" int pirg_server_save_to_file(pirg_server* S, const char* file_path, char* err,
                                        size_t errlen) {
  if (S == nullptr || file_path == nullptr) {
    return Report(Invalid("server and file_path must not be null."), err, errlen);
  }
  FILE* f = std::fopen(file_path, "wb");
  if (f == nullptr) {
    return Report(Error{PIRG_INTERNAL, std::string("Failed to open file for writing: ") +
                                           file_path},
                  err, errlen);
  }
  ImageSink out(f);
  Error er = SaveImage(S, out, file_path);
  const bool closed = std::fclose(f) == 0;
  if (er.ok() && !closed) er = WriteFailed(file_path);
  return er.ok() ? 0 : Report(er, err, errlen);
}

extern "C" int pirg_server_save_to_buffer(pirg_server* S, uint8_t* buffer, size_t capacity,
                                          size_t* size, char* err, size_t errlen) {
  if (S == nullptr || size == nullptr) {
    return Report(Invalid("server and size must not be null."), err, errlen);
  }
  const uint64_t need = ImageBytes(S);
  *size = static_cast<size_t>(need);
  if (buffer == nullptr) return 0;
  if (capacity < need) {
    return Report(Invalid("buffer holds " + std::to_string(capacity) + " bytes; the image has " +
                          std::to_string(need) + "."),
                  err, errlen);
  }
  ImageSink out(buffer, capacity);
  Error er = SaveImage(S, out, "buffer");
  if (er.ok() && out.written() != need) er = Internal("image size mismatch");
  return er.ok() ? 0 : Report(er, err, errlen);
}

extern "C" int pirg_server_load_from_file(const pirg_params* params, const char* file_path,
                                          pirg_server** out, char* err, size_t errlen) {
  if (params == nullptr || file_path == nullptr || out == nullptr) {
    return Report(Invalid("params, file_path and out must not be null."), err, errlen);
  }
  // file_util::GetFileContents's failure (pir_server.cc:799-801); the file is
  // mapped rather than copied into memory.
  const Error read_failed{PIRG_INTERNAL, "Failed to read file."};
  const int fd = ::open(file_path, O_RDONLY);
  if (fd < 0) return Report(read_failed, err, errlen);
  struct stat st {};
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    return Report(read_failed, err, errlen);
  }
  const size_t size = static_cast<size_t>(st.st_size);
  void* map = nullptr;
  if (size > 0) {
    map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (map == MAP_FAILED) {
      ::close(fd);
      return Report(read_failed, err, errlen);
    }
  }
  ::close(fd);
  std::unique_ptr<pirg_server> S;
  Error er = LoadFromBuffer(params, static_cast<const uint8_t*>(map), size, &S);
  if (map != nullptr) ::munmap(map, size);
  if (!er.ok()) return Report(er, err, errlen);
  *out = S.release();
  return 0;
}

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
