// Preprocessed material in the reference's protobuf forms, straight to a
// resident server:
//   * proto::PirPreprocessedData (crypto/proto/pir_server.proto:54-66), the
//     file of SavePirPreprocessedDataToFile (pir_preprocessed_data.h:150-160)
//     read as DeserializePirPreprocessedData does (:93-148);
//   * the directory of SavePirPreprocessedDataToDirectory (:175-224):
//     metadata.bin (PirPreprocessedMetadata) and shard_<i>.bin
//     (PirPreprocessedShard), read as LoadPirPreprocessedDataFromDirectory
//     does (:226-320).
// Messages are walked with a minimal protobuf wire-format reader (varint and
// length-delimited fields; unknown fields skipped). Matrices are taken as the
// raw bytes Matrix::ToProto wrote (crypto/matrix.cc:479-495) with
// Matrix::FromProto's checks (:497-532); polynomials are decoded on the GPU
// with CreateFromProto's checks (polynomial.cc:208-252). The result goes
// through PirServer::Create's validation and upload.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pirgpu.h"
#include "server_internal.h"

namespace pirg {
namespace {

struct Span {
  const uint8_t* p = nullptr;
  size_t n = 0;
};

// Protobuf wire format: tag = field << 3 | wire type; 0 varint, 1 fixed64,
// 2 length-delimited, 5 fixed32.
class PbReader {
 public:
  explicit PbReader(Span s) : p_(s.p), end_(s.p + s.n) {}
  bool ok() const { return ok_; }
  bool done() const { return p_ >= end_; }
  uint64_t Varint() {
    uint64_t v = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p_ >= end_) {
        ok_ = false;
        return 0;
      }
      const uint8_t b = *p_++;
      v |= static_cast<uint64_t>(b & 0x7f) << shift;
      if ((b & 0x80) == 0) return v;
    }
    ok_ = false;
    return 0;
  }
  bool Next(int* field, int* wire) {
    if (done() || !ok_) return false;
    const uint64_t tag = Varint();
    *field = static_cast<int>(tag >> 3);
    *wire = static_cast<int>(tag & 7);
    if (*field == 0) ok_ = false;
    return ok_;
  }
  Span Bytes() {
    const uint64_t n = Varint();
    if (!ok_ || n > static_cast<uint64_t>(end_ - p_)) {
      ok_ = false;
      return {};
    }
    Span s{p_, static_cast<size_t>(n)};
    p_ += n;
    return s;
  }
  void Skip(int wire) {
    switch (wire) {
      case 0: (void)Varint(); break;
      case 1: Advance(8); break;
      case 2: (void)Bytes(); break;
      case 5: Advance(4); break;
      default: ok_ = false;
    }
  }

 private:
  void Advance(size_t n) {
    if (n > static_cast<size_t>(end_ - p_)) {
      ok_ = false;
      return;
    }
    p_ += n;
  }
  const uint8_t* p_;
  const uint8_t* end_;
  bool ok_ = true;
};

struct MatrixRec {
  Span data, condensed;
  bool has_data = false, has_condensed = false;
  int rows = 0, cols = 0, original_cols = 0;
  bool is_condensed = false;
};

struct OutputRec {
  Span a_tilde;
  MatrixRec matrix;
  std::vector<Span> t_vec_h;
};

struct Material {
  std::vector<MatrixRec> db;
  std::vector<std::vector<OutputRec>> outputs;
  std::vector<Span> second_dim_a;
};

bool ParsePolynomial(Span s, Span* out) {
  PbReader r(s);
  int f, w;
  *out = Span{};
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      *out = r.Bytes();
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

bool ParseMatrix(Span s, MatrixRec* m) {
  PbReader r(s);
  int f, w;
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      m->data = r.Bytes();
      m->has_data = true;
      m->has_condensed = false;  // oneof: the last one set wins
    } else if (f == 2 && w == 2) {
      m->condensed = r.Bytes();
      m->has_condensed = true;
      m->has_data = false;
    } else if (f == 3 && w == 0) {
      m->rows = static_cast<int>(r.Varint());
    } else if (f == 4 && w == 0) {
      m->cols = static_cast<int>(r.Varint());
    } else if (f == 5 && w == 0) {
      m->is_condensed = r.Varint() != 0;
    } else if (f == 6 && w == 0) {
      m->original_cols = static_cast<int>(r.Varint());
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

bool ParseOutput(Span s, OutputRec* o) {
  PbReader r(s);
  int f, w;
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      if (!ParsePolynomial(r.Bytes(), &o->a_tilde)) return false;
    } else if (f == 2 && w == 2) {
      if (!ParseMatrix(r.Bytes(), &o->matrix)) return false;
    } else if (f == 3 && w == 2) {
      Span p;
      if (!ParsePolynomial(r.Bytes(), &p)) return false;
      o->t_vec_h.push_back(p);
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

bool ParseShardOutputs(Span s, std::vector<OutputRec>* outs) {
  PbReader r(s);
  int f, w;
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      outs->emplace_back();
      if (!ParseOutput(r.Bytes(), &outs->back())) return false;
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

// PirPreprocessedData: 1 db_matrices, 2 preprocessed_outputs, 3 second_dim_a.
bool ParseData(Span s, Material* m) {
  PbReader r(s);
  int f, w;
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      m->db.emplace_back();
      if (!ParseMatrix(r.Bytes(), &m->db.back())) return false;
    } else if (f == 2 && w == 2) {
      m->outputs.emplace_back();
      if (!ParseShardOutputs(r.Bytes(), &m->outputs.back())) return false;
    } else if (f == 3 && w == 2) {
      Span p;
      if (!ParsePolynomial(r.Bytes(), &p)) return false;
      m->second_dim_a.push_back(p);
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

// PirPreprocessedShard: 1 db_matrix, 2 outputs.
bool ParseShard(Span s, Material* m) {
  PbReader r(s);
  int f, w;
  m->db.emplace_back();
  m->outputs.emplace_back();
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      if (!ParseMatrix(r.Bytes(), &m->db.back())) return false;
    } else if (f == 2 && w == 2) {
      m->outputs.back().emplace_back();
      if (!ParseOutput(r.Bytes(), &m->outputs.back().back())) return false;
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

// PirPreprocessedMetadata: 1 second_dim_a.
bool ParseMetadata(Span s, Material* m) {
  PbReader r(s);
  int f, w;
  while (r.Next(&f, &w)) {
    if (f == 1 && w == 2) {
      Span p;
      if (!ParsePolynomial(r.Bytes(), &p)) return false;
      m->second_dim_a.push_back(p);
    } else {
      r.Skip(w);
    }
  }
  return r.ok();
}

// Matrix::FromProto's checks (crypto/matrix.cc:500-527); returns the element
// bytes.
Error MatrixBytes(const MatrixRec& m, size_t elem_size, Span* out) {
  if (m.is_condensed && m.has_data) {
    return Invalid("Condensed matrix cannot have normal data set.");
  }
  if (!m.is_condensed && m.has_condensed) {
    return Invalid("Normal matrix cannot have condensed data set.");
  }
  if (m.has_data && m.data.n % elem_size != 0) return Invalid("Invalid serialized data size.");
  if (m.has_condensed && m.condensed.n % 8 != 0) {
    return Invalid("Invalid serialized condensed_data size.");
  }
  *out = m.is_condensed ? m.condensed : m.data;
  return Ok();
}

// Logical columns of a matrix record (Matrix::Cols() is original_cols).
int LogicalCols(const MatrixRec& m) { return m.original_cols; }

Error Build(const pirg_params* params, const Material& mat, std::unique_ptr<pirg_server>* out) {
  pirg_server probe;
  Error er = DeriveServer(*params, &probe);
  if (!er.ok()) return er;
  const int d = probe.d, t = probe.t, ell = probe.ell, k = probe.k;
  const int logm = params->log_q >= 64 ? 64 : params->log_q;
  // Matrices: checks, then the byte counts the layouts need.
  std::vector<const uint64_t*> db_ptrs;
  std::vector<int> db_rows, db_cols;
  for (const MatrixRec& m : mat.db) {
    Span b;
    er = MatrixBytes(m, params->db_bytes, &b);
    if (!er.ok()) return er;
    db_rows.push_back(m.rows);
    db_cols.push_back(LogicalCols(m));
    const size_t need = static_cast<size_t>(m.rows) * LogicalCols(m) * params->db_bytes;
    if (b.n < need) return Invalid("DB matrix bytes are short of rows x columns.");
    db_ptrs.push_back(reinterpret_cast<const uint64_t*>(b.p));
  }
  std::vector<int> olen;
  std::vector<const void*> pack_ptrs;
  std::vector<Span> polys;  // a_tilde (per output), then t_vec_h, then second_dim_a
  std::vector<const OutputRec*> outs;
  for (const auto& shard : mat.outputs) {
    olen.push_back(static_cast<int>(shard.size()));
    for (const OutputRec& o : shard) outs.push_back(&o);
  }
  for (const OutputRec* o : outs) {
    Span b;
    er = MatrixBytes(o->matrix, params->mat_bytes, &b);
    if (!er.ok()) return er;
    const size_t need = static_cast<size_t>(d) * k * d * params->mat_bytes;
    if (o->matrix.rows != d || LogicalCols(o->matrix) != k * d || b.n < need) {
      return Invalid("pack matrix must be d x k*d.");
    }
    if (o->t_vec_h.size() != static_cast<size_t>(k)) {
      return Invalid("t_vec_h must hold k polynomials.");
    }
    pack_ptrs.push_back(b.p);
  }
  for (const OutputRec* o : outs) polys.push_back(o->a_tilde);
  for (const OutputRec* o : outs) {
    for (const Span& p : o->t_vec_h) polys.push_back(p);
  }
  for (const Span& p : mat.second_dim_a) polys.push_back(p);
  // CreateFromProto's checks (polynomial.cc:211-226) for each polynomial.
  const size_t poly_bytes = (static_cast<size_t>(d) * logm + 7) / 8;
  for (const Span& p : polys) {
    if (p.n == 0) return Invalid("Empty serialized polynomial.");
    if (p.n != poly_bytes) return Invalid("Unexpected serialized size.");
  }
  er = CheckDevice();
  if (!er.ok()) return er;
  // Decode every polynomial on the GPU.
  const long long pitch = StreamWords(d, logm);
  std::vector<u64> staged(polys.size() * static_cast<size_t>(pitch), 0);
  for (size_t i = 0; i < polys.size(); ++i) {
    std::memcpy(staged.data() + i * pitch, polys[i].p, polys[i].n);
  }
  DevBuf words, dev;
  if (!polys.empty()) {
    er = words.alloc(staged.size() * 8);
    if (!er.ok()) return er;
    er = dev.alloc(polys.size() * static_cast<size_t>(d) * 8);
    if (!er.ok()) return er;
    CallStream cs;
    PIRG_HIP(cs.create());
    PIRG_HIP(CopySync(words.p, staged.data(), staged.size() * 8, hipMemcpyHostToDevice, cs.s));
    PIRG_HIP(LaunchBitUnpack(words.as<u64>(), pitch, 0, static_cast<int>(polys.size()), d, logm,
                             false, 8, dev.p, cs.s));
    PIRG_HIP(hipStreamSynchronize(cs.s));
  }
  const size_t no = outs.size();
  const u64* base = dev.as<u64>();
  std::vector<const uint64_t*> at(no), th(no);
  for (size_t i = 0; i < no; ++i) {
    at[i] = base + i * d;
    th[i] = base + (no + i * k) * static_cast<size_t>(d);
  }
  const u64* sda = mat.second_dim_a.empty() ? nullptr : base + (no + no * k) * d;
  pirg_preprocessed_view v{};
  v.num_db_matrices = static_cast<int>(mat.db.size());
  v.db_cdata = db_ptrs.data();
  v.db_rows = db_rows.data();
  v.db_cols = db_cols.data();
  v.num_outputs = static_cast<int>(mat.outputs.size());
  v.outputs_len = olen.data();
  v.pack = pack_ptrs.data();
  v.a_tilde = at.data();
  v.t_vec_h = th.data();
  v.second_dim_a_len = static_cast<int>(mat.second_dim_a.size());
  v.second_dim_a = sda;
  (void)t;
  (void)ell;
  return CreateFromView(params, &v, out);
}

Error ReadFile(const std::string& path, std::string* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) return Error{5, "Failed to open file: " + path};  // NOT_FOUND
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out->resize(n > 0 ? static_cast<size_t>(n) : 0);
  const size_t got = n > 0 ? std::fread(&(*out)[0], 1, out->size(), f) : 0;
  std::fclose(f);
  if (got != out->size()) return Error{13, "Failed to read file: " + path};
  return Ok();
}

}  // namespace
}  // namespace pirg

using namespace pirg;

extern "C" {

int pirg_server_create_from_proto(const pirg_params* params, const uint8_t* bytes, size_t size,
                                  pirg_server** out, char* err, size_t errlen) {
  if (params == nullptr || out == nullptr || (bytes == nullptr && size > 0)) {
    return Report(Invalid("params, bytes and out must not be null."), err, errlen);
  }
  Material mat;
  if (!ParseData(Span{bytes, size}, &mat)) {
    return Report(Invalid("Failed to parse PirPreprocessedData from file."), err, errlen);
  }
  std::unique_ptr<pirg_server> S;
  Error er = Build(params, mat, &S);
  if (!er.ok()) return Report(er, err, errlen);
  *out = S.release();
  return 0;
}

int pirg_server_create_from_directory(const pirg_params* params, const char* dir_path,
                                      pirg_server** out, char* err, size_t errlen) {
  if (params == nullptr || out == nullptr || dir_path == nullptr) {
    return Report(Invalid("params, dir_path and out must not be null."), err, errlen);
  }
  const std::string dir(dir_path);
  std::string meta;
  Error er = ReadFile(dir + "/metadata.bin", &meta);
  if (!er.ok()) return Report(er, err, errlen);
  Material mat;
  if (!ParseMetadata(Span{reinterpret_cast<const uint8_t*>(meta.data()), meta.size()}, &mat)) {
    return Report(Invalid("Failed to parse PirPreprocessedMetadata."), err, errlen);
  }
  const int ell = params->entry_size_multiple;
  std::vector<std::string> shards(ell > 0 ? ell : 0);
  for (int i = 0; i < ell; ++i) {
    er = ReadFile(dir + "/shard_" + std::to_string(i) + ".bin", &shards[i]);
    if (!er.ok()) return Report(er, err, errlen);
    if (!ParseShard(Span{reinterpret_cast<const uint8_t*>(shards[i].data()), shards[i].size()},
                    &mat)) {
      return Report(Invalid("Failed to parse PirPreprocessedShard " + std::to_string(i)), err,
                    errlen);
    }
  }
  std::unique_ptr<pirg_server> S;
  er = Build(params, mat, &S);
  if (!er.ok()) return Report(er, err, errlen);
  *out = S.release();
  return 0;
}

}  // extern "C"
