// Truncated and mutated inputs through the host-side parsers of libpirgpu:
// the server image of PirServer::SaveToFile (pirg_server_load_from_buffer,
// crypto/pir/server/pir_server.cc:594-792) and a serialized
// PirPreprocessedData message (pirg_server_create_from_proto,
// crypto/pir/pir_preprocessed_data.h:93-148).
//
// Built with AddressSanitizer on the host code only and run on a machine
// without a GPU: an input that parses stops at the device check
// (PIRG_UNAVAILABLE), so only the parsers run, and every other input must be
// refused with an error code. Any read outside an input ends the run with
// AddressSanitizer's report and a non-zero status.
//
// usage: parser_fuzz params.bin image.bin proto.bin seed iterations
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pirgpu.h"

namespace {

std::vector<uint8_t> ReadAll(const char* path) {
  std::vector<uint8_t> out;
  FILE* f = std::fopen(path, "rb");
  if (f == nullptr) return out;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
  std::fclose(f);
  return out;
}

struct Rng {
  uint64_t s;
  uint64_t Next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  size_t Below(size_t n) { return n == 0 ? 0 : static_cast<size_t>(Next() % n); }
};

using Entry = int (*)(const pirg_params*, const uint8_t*, size_t, pirg_server**, char*, size_t);

struct Runner {
  const char* name;
  const pirg_params* P;
  Entry fn;
  std::map<int, long> codes;
  long calls = 0, created = 0;

  // The input in a heap block of exactly its size, so a read past its end
  // is one AddressSanitizer sees.
  void Try(const std::vector<uint8_t>& in) {
    std::vector<uint8_t> exact(in.begin(), in.end());
    exact.shrink_to_fit();
    pirg_server* out = nullptr;
    char err[512];
    err[0] = 0;
    const int rc = fn(P, exact.empty() ? nullptr : exact.data(), exact.size(), &out, err,
                      sizeof err);
    ++calls;
    ++codes[rc];
    if (rc == 0) {
      ++created;
      pirg_server_destroy(out);
    }
  }

  void Report() const {
    std::printf("%s: %ld calls", name, calls);
    for (const auto& kv : codes) std::printf(" rc%d=%ld", kv.first, kv.second);
    std::printf(" created=%ld\n", created);
  }
};

const uint64_t kExtremes[] = {0,           1,          0x7f,       0x80,       0xff,
                              0x7fffffff,  0x80000000, 0xffffffff, 1ULL << 32, 1ULL << 62,
                              1ULL << 63,  ~0ULL};

// Truncations (every prefix over the first `dense` bytes, then a spread) and
// mutations: single bytes, runs of bytes and little-endian 4 / 8-byte fields
// set to extreme values, at positions drawn from the given hot regions half
// the time and from the whole input otherwise.
void Fuzz(Runner& r, const std::vector<uint8_t>& base, const std::vector<size_t>& hot,
          Rng& rng, long iters) {
  r.Try(base);
  const size_t dense = std::min<size_t>(base.size(), 4096);
  for (size_t n = 0; n < dense; ++n) r.Try(std::vector<uint8_t>(base.begin(), base.begin() + n));
  for (int i = 0; i < 512; ++i) {
    const size_t n = dense + rng.Below(base.size() - dense + 1);
    r.Try(std::vector<uint8_t>(base.begin(), base.begin() + std::min(n, base.size())));
  }
  for (long it = 0; it < iters; ++it) {
    std::vector<uint8_t> m = base;
    const int edits = 1 + static_cast<int>(rng.Below(4));
    for (int e = 0; e < edits; ++e) {
      size_t pos;
      if (!hot.empty() && rng.Below(2) == 0) {
        pos = hot[rng.Below(hot.size())] + rng.Below(32);
      } else {
        pos = rng.Below(m.size());
      }
      if (pos >= m.size()) pos = m.size() - 1;
      switch (rng.Below(4)) {
        case 0:
          m[pos] = static_cast<uint8_t>(rng.Next());
          break;
        case 1: {
          const size_t len = 1 + rng.Below(16);
          for (size_t j = pos; j < std::min(m.size(), pos + len); ++j) m[j] = 0xff;
          break;
        }
        default: {
          const uint64_t v = kExtremes[rng.Below(sizeof kExtremes / sizeof kExtremes[0])];
          const size_t w = rng.Below(2) == 0 ? 4 : 8;
          for (size_t j = 0; j < w && pos + j < m.size(); ++j) {
            m[pos + j] = static_cast<uint8_t>(v >> (8 * j));
          }
        }
      }
    }
    if (rng.Below(8) == 0) m.resize(rng.Below(m.size() + 1));
    r.Try(m);
  }
}

uint64_t U64At(const std::vector<uint8_t>& b, size_t off) {
  uint64_t v = 0;
  if (off + 8 <= b.size()) std::memcpy(&v, b.data() + off, 8);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s params.bin image.bin proto.bin seed iterations\n", argv[0]);
    return 2;
  }
  const std::vector<uint8_t> pbytes = ReadAll(argv[1]);
  if (pbytes.size() != sizeof(pirg_params)) {
    std::fprintf(stderr, "params.bin holds %zu bytes, pirg_params %zu\n", pbytes.size(),
                 sizeof(pirg_params));
    return 2;
  }
  pirg_params P;
  std::memcpy(&P, pbytes.data(), sizeof P);
  const std::vector<uint8_t> image = ReadAll(argv[2]);
  const std::vector<uint8_t> proto = ReadAll(argv[3]);
  if (image.empty() || proto.empty()) {
    std::fprintf(stderr, "empty image or proto input\n");
    return 2;
  }
  Rng rng{std::stoull(argv[4]) * 0x9E3779B97F4A7C15ULL + 1};
  const long iters = std::stol(argv[5]);

  // The image's header fields: the DB header at 0, the pack header after the
  // DB words (its element count at byte 13), the outputs after the pack
  // bytes (the pack byte count at pack header + 21), and the tail.
  std::vector<size_t> hot = {0};
  const uint64_t db_elems = U64At(image, 13);
  const size_t pk_hdr = 21 + static_cast<size_t>(db_elems) * 8;
  if (pk_hdr < image.size()) {
    hot.push_back(pk_hdr);
    const size_t out_hdr = pk_hdr + 29 + static_cast<size_t>(U64At(image, pk_hdr + 21));
    if (out_hdr < image.size()) hot.push_back(out_hdr);
  }
  for (int i = 1; i <= 16; ++i) hot.push_back(image.size() - image.size() * i / 64);
  Runner img{"load_from_buffer", &P, &pirg_server_load_from_buffer};
  Fuzz(img, image, hot, rng, iters);

  std::vector<size_t> phot = {0};
  for (int i = 1; i <= 16; ++i) phot.push_back(proto.size() - proto.size() * i / 64);
  Runner pb{"create_from_proto", &P, &pirg_server_create_from_proto};
  Fuzz(pb, proto, phot, rng, iters);

  img.Report();
  pb.Report();
  // The intact inputs parse (stopping at the device check) and the rest are
  // refused: nothing is created without a GPU.
  return (img.created == 0 && pb.created == 0) ? 0 : 1;
}
