// mh_io.cpp -- lattice-file readers and writers (host code): MILC's version-5
// binary gauge format and the SciDAC / ILDG formats on LIME records, for gauge
// links, colour vectors and colour-matrix fields.
//
// Formats (kept byte-compatible with the files MILC and QIO write):
//   v5 gauge     generic/io_lat_utils.c:370-389: magic 20103, dims[4],
//                time_stamp[64], order (88 bytes), [site list of V int32 when
//                order != 0], checksum sum29, sum31 (io_lat_utils.c:766-812,
//                io_lat4.c:570-585), then per site 4 x fsu3_matrix (float).
//                Byte order from the magic number (io_lat_utils.c:1305-1345).
//                Checksum: every data word in native order, rotated left by
//                its word rank mod 29 / mod 31 and XORed (io_lat4.c:694-707).
//   LIME         144-byte big-endian record header (magic 0x456789ab,
//                include/file_types.h:10; version; flags; payload length;
//                128-byte type), payload padded to a multiple of 8 bytes.
//   SciDAC/ILDG  binary payload big-endian, lexicographic sites (x fastest),
//                datacount objects per site; the checksum record holds QIO's
//                DML checksum: per site c = CRC-32 of the site's bytes as
//                stored, suma ^= rotl(c, rank % 29), sumb ^= rotl(c, rank % 31).
//                QIO itself is an external library (not in the reference
//                tree); this follows its published algorithm.
//
// The checksum, byte swap, precision change and reorder to MILC's node_index
// order (even sites first) run in one pass over the payload, split over host
// threads (XOR is order-free, so per-thread partial checksums combine exactly).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/milc_io.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the readers assume a little-endian host");

namespace {

const uint32_t kV5Magic = 20103;
const uint32_t kLimeMagic = 0x456789abu;

// CRC-32 (IEEE 802.3: reflected polynomial 0xEDB88320, register and result
// inverted), eight table lookups per 8 input bytes.
struct Crc32 {
  uint32_t t[8][256];
  Crc32() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; i++)
      for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xffu];
  }
  uint32_t operator()(const uint8_t* p, size_t n) const {
    uint32_t c = 0xFFFFFFFFu;
    while (n >= 8) {
      uint32_t lo, hi;
      memcpy(&lo, p, 4);
      memcpy(&hi, p + 4, 4);
      lo ^= c;
      c = t[7][lo & 0xffu] ^ t[6][(lo >> 8) & 0xffu] ^ t[5][(lo >> 16) & 0xffu] ^ t[4][lo >> 24] ^
          t[3][hi & 0xffu] ^ t[2][(hi >> 8) & 0xffu] ^ t[1][(hi >> 16) & 0xffu] ^ t[0][hi >> 24];
      p += 8;
      n -= 8;
    }
    while (n--) c = t[0][(c ^ *p++) & 0xffu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
  }
};
const Crc32& crc32() {
  static const Crc32 k;
  return k;
}

inline uint32_t rotl(uint32_t w, unsigned r) { return r ? (w << r) | (w >> (32 - r)) : w; }
inline uint32_t ld32(const uint8_t* p, bool swap) {
  uint32_t w;
  memcpy(&w, p, 4);
  return swap ? __builtin_bswap32(w) : w;
}
inline uint64_t ld64be(const uint8_t* p) {
  uint64_t w;
  memcpy(&w, p, 8);
  return __builtin_bswap64(w);
}

// Runs f(begin, end, thread) over [0, n) on up to 16 host threads.
template <class F>
void parallel_for(size_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  if (nt == 1) {
    f(size_t(0), n, 0u);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; k++) th.emplace_back(f, n * k / nt, n * (k + 1) / nt, k);
  for (auto& t : th) t.join();
}

struct Lattice {
  int d[4];
  size_t V;
  explicit Lattice(const int* dims) {
    V = 1;
    for (int k = 0; k < 4; k++) {
      d[k] = dims[k];
      V *= (size_t)dims[k];
    }
  }
  // MILC node_index of lexicographic site l (single rank, even sites first)
  size_t node_index(size_t l) const {
    size_t x = l % d[0], r = l / d[0];
    size_t y = r % d[1];
    r /= d[1];
    size_t z = r % d[2], t = r / d[2];
    return ((x + y + z + t) & 1) ? (l + V) / 2 : l / 2;
  }
};

bool dims_ok(const int* d) {
  if (!d) return false;
  size_t V = 1;
  for (int k = 0; k < 4; k++) {
    if (d[k] <= 0 || d[k] > (1 << 16)) return false;
    V *= (size_t)d[k];
  }
  // the node_index order needs an even volume split into two equal parities
  return V % 2 == 0 && V <= (size_t(1) << 34);
}

struct FileBuf {
  std::unique_ptr<uint8_t[]> p;
  size_t n = 0;
};
int read_file(const char* path, FileBuf& b) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  if (fseeko(f, 0, SEEK_END) != 0) {
    fclose(f);
    return -1;
  }
  off_t len = ftello(f);
  if (len < 0 || fseeko(f, 0, SEEK_SET) != 0) {
    fclose(f);
    return -1;
  }
  b.n = (size_t)len;
  b.p.reset(new (std::nothrow) uint8_t[b.n ? b.n : 1]);
  if (!b.p) {
    fclose(f);
    return -1;
  }
  size_t got = b.n ? fread(b.p.get(), 1, b.n, f) : 0;
  fclose(f);
  return got == b.n ? 0 : -1;
}

void set_info(milc_io_info* info, const milc_io_info& v) {
  if (info) *info = v;
}

// ---------------------------------------------------------------------------
// MILC version-5 gauge files
// ---------------------------------------------------------------------------
int read_v5(const uint8_t* d, size_t n, const int* want, double* links, milc_io_info* info) {
  if (n < 88) return -1;
  uint32_t m;
  memcpy(&m, d, 4);
  const bool swap = m != kV5Magic;
  if (swap && __builtin_bswap32(m) != kV5Magic) return -2;
  int dims[4];
  for (int k = 0; k < 4; k++) dims[k] = (int)ld32(d + 4 + 4 * k, swap);
  const int order = (int)ld32(d + 84, swap);
  if (!dims_ok(dims)) return -2;
  if (want)
    for (int k = 0; k < 4; k++)
      if (dims[k] != want[k]) return -4;
  const Lattice L(dims);
  size_t off = 88;
  const uint8_t* list = nullptr;
  if (order != 0) {
    if (n < off + 4 * L.V) return -1;
    list = d + off;
    off += 4 * L.V;
    for (size_t s = 0; s < L.V; s++)
      if ((size_t)ld32(list + 4 * s, swap) >= L.V) return -2;
  }
  const size_t site_bytes = 4 * 18 * 4;
  if (n < off + 8 + L.V * site_bytes) return -1;
  milc_io_info inf{};
  inf.format = MILC_IO_V5;
  inf.precision = 1;
  for (int k = 0; k < 4; k++) inf.dims[k] = dims[k];
  inf.count = 4;
  inf.datatype = 1;
  inf.byte_reversed = swap;
  inf.has_checksum = 1;
  inf.stored_a = ld32(d + off, swap);
  inf.stored_b = ld32(d + off + 4, swap);
  const uint8_t* data = d + off + 8;
  std::vector<uint32_t> p29(16, 0), p31(16, 0);
  parallel_for(L.V, [&](size_t b, size_t e, unsigned tid) {
    uint32_t a29 = 0, a31 = 0;
    size_t k = b * 72;  // word rank of the first word of site b
    unsigned r29 = (unsigned)(k % 29), r31 = (unsigned)(k % 31);
    for (size_t s = b; s < e; s++) {
      const uint8_t* src = data + s * site_bytes;
      double* dst = nullptr;
      if (links) {
        size_t l = list ? (size_t)ld32(list + 4 * s, swap) : s;
        dst = links + L.node_index(l) * 72;
      }
      for (int j = 0; j < 72; j++) {
        uint32_t w = ld32(src + 4 * j, swap);
        a29 ^= rotl(w, r29);
        a31 ^= rotl(w, r31);
        if (++r29 == 29) r29 = 0;
        if (++r31 == 31) r31 = 0;
        if (dst) {
          float f;
          memcpy(&f, &w, 4);
          dst[j] = (double)f;
        }
      }
    }
    p29[tid] = a29;
    p31[tid] = a31;
  });
  for (int t = 0; t < 16; t++) {
    inf.computed_a ^= p29[t];
    inf.computed_b ^= p31[t];
  }
  inf.checksum_ok = inf.computed_a == inf.stored_a && inf.computed_b == inf.stored_b;
  set_info(info, inf);
  return 0;
}

int write_v5(const char* path, const int* dims, const double* links, milc_io_info* info) {
  const Lattice L(dims);
  const size_t site_bytes = 4 * 18 * 4;
  std::unique_ptr<uint8_t[]> buf(new (std::nothrow) uint8_t[88 + 8 + L.V * site_bytes]);
  if (!buf) return -1;
  uint8_t* h = buf.get();
  memset(h, 0, 88);
  uint32_t w = kV5Magic;
  memcpy(h, &w, 4);
  for (int k = 0; k < 4; k++) {
    int32_t v = dims[k];
    memcpy(h + 4 + 4 * k, &v, 4);
  }
  time_t now = time(nullptr);
  char stamp[64] = {0};
  struct tm tmv;
  gmtime_r(&now, &tmv);
  strftime(stamp, sizeof stamp, "%a %b %e %H:%M:%S %Y", &tmv);  // ctime's layout
  memcpy(h + 20, stamp, 64);  // order (h + 84) stays 0: natural site order
  uint8_t* data = h + 96;
  std::vector<uint32_t> p29(16, 0), p31(16, 0);
  parallel_for(L.V, [&](size_t b, size_t e, unsigned tid) {
    uint32_t a29 = 0, a31 = 0;
    size_t k = b * 72;
    unsigned r29 = (unsigned)(k % 29), r31 = (unsigned)(k % 31);
    for (size_t s = b; s < e; s++) {
      const double* src = links + L.node_index(s) * 72;
      uint8_t* dst = data + s * site_bytes;
      for (int j = 0; j < 72; j++) {
        float f = (float)src[j];
        uint32_t u;
        memcpy(&u, &f, 4);
        memcpy(dst + 4 * j, &u, 4);
        a29 ^= rotl(u, r29);
        a31 ^= rotl(u, r31);
        if (++r29 == 29) r29 = 0;
        if (++r31 == 31) r31 = 0;
      }
    }
    p29[tid] = a29;
    p31[tid] = a31;
  });
  uint32_t s29 = 0, s31 = 0;
  for (int t = 0; t < 16; t++) {
    s29 ^= p29[t];
    s31 ^= p31[t];
  }
  memcpy(h + 88, &s29, 4);
  memcpy(h + 92, &s31, 4);
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  size_t total = 96 + L.V * site_bytes;
  bool ok = fwrite(h, 1, total, f) == total;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return -1;
  milc_io_info inf{};
  inf.format = MILC_IO_V5;
  inf.precision = 1;
  for (int k = 0; k < 4; k++) inf.dims[k] = dims[k];
  inf.count = 4;
  inf.datatype = 1;
  inf.has_checksum = 1;
  inf.checksum_ok = 1;
  inf.stored_a = inf.computed_a = s29;
  inf.stored_b = inf.computed_b = s31;
  set_info(info, inf);
  return 0;
}

// ---------------------------------------------------------------------------
// LIME, SciDAC and ILDG
// ---------------------------------------------------------------------------
struct Rec {
  std::string type;
  size_t off, len;
};

int lime_scan(const uint8_t* d, size_t n, std::vector<Rec>& recs) {
  size_t off = 0;
  while (off < n) {
    if (n - off < 144) return -1;
    uint32_t m;
    memcpy(&m, d + off, 4);
    if (__builtin_bswap32(m) != kLimeMagic) return -2;
    uint64_t len = ld64be(d + off + 8);
    char typ[129];
    memcpy(typ, d + off + 16, 128);
    typ[128] = 0;
    size_t padded = (size_t)((len + 7) / 8 * 8);
    if (len > n || n - off - 144 < (size_t)len) return -1;
    recs.push_back(Rec{std::string(typ), off + 144, (size_t)len});
    off += 144 + std::min(padded, n - off - 144);
  }
  return recs.empty() ? -2 : 0;
}

const Rec* find_rec(const std::vector<Rec>& recs, const char* type) {
  for (const Rec& r : recs)
    if (r.type == type) return &r;
  return nullptr;
}

// text between <tag> and </tag> in a record payload ("" when absent)
std::string xml_tag(const uint8_t* d, const Rec* r, const char* tag) {
  if (!r) return "";
  std::string s((const char*)d + r->off, r->len);
  std::string open = std::string("<") + tag + ">", close = std::string("</") + tag + ">";
  size_t a = s.find(open);
  if (a == std::string::npos) return "";
  a += open.size();
  size_t b = s.find(close, a);
  if (b == std::string::npos) return "";
  return s.substr(a, b - a);
}

struct LimeField {
  int format = 0, dims[4] = {0, 0, 0, 0}, datatype = 0, count = 0, precision = 0;
  const Rec* bin = nullptr;
  const Rec* cks = nullptr;
};

int lime_describe(const uint8_t* d, const std::vector<Rec>& recs, LimeField& f) {
  const Rec* ildg = find_rec(recs, "ildg-binary-data");
  if (ildg) {
    const Rec* fmt = find_rec(recs, "ildg-format");
    if (!fmt) return -2;
    const char* t[4] = {"lx", "ly", "lz", "lt"};
    for (int k = 0; k < 4; k++) f.dims[k] = atoi(xml_tag(d, fmt, t[k]).c_str());
    int p = atoi(xml_tag(d, fmt, "precision").c_str());
    if (p != 32 && p != 64) return -2;
    f.format = MILC_IO_ILDG;
    f.precision = p == 32 ? 1 : 2;
    f.datatype = 1;
    f.count = 4;
    f.bin = ildg;
  } else {
    const Rec* pf = find_rec(recs, "scidac-private-file-xml");
    const Rec* pr = find_rec(recs, "scidac-private-record-xml");
    f.bin = find_rec(recs, "scidac-binary-data");
    if (!pf || !pr || !f.bin) return -2;
    std::string ds = xml_tag(d, pf, "dims");
    if (sscanf(ds.c_str(), "%d %d %d %d", &f.dims[0], &f.dims[1], &f.dims[2], &f.dims[3]) != 4) return -2;
    std::string dt = xml_tag(d, pr, "datatype"), pc = xml_tag(d, pr, "precision");
    f.count = atoi(xml_tag(d, pr, "datacount").c_str());
    int typesize = atoi(xml_tag(d, pr, "typesize").c_str());
    f.precision = pc == "F" ? 1 : (pc == "D" ? 2 : 0);
    if (dt.size() > 11 && dt.compare(dt.size() - 11, 11, "ColorMatrix") == 0) f.datatype = 1;
    else if (dt.size() > 11 && dt.compare(dt.size() - 11, 11, "ColorVector") == 0) f.datatype = 2;
    if (!f.precision || !f.datatype || f.count <= 0 || f.count > 64) return -2;
    if (typesize != (f.datatype == 1 ? 18 : 6) * (f.precision == 1 ? 4 : 8)) return -2;
    f.format = MILC_IO_SCIDAC;
  }
  f.cks = find_rec(recs, "scidac-checksum");
  if (!dims_ok(f.dims)) return -2;
  return 0;
}

// reads the field, checksums it and (out != nullptr) stores it in node_index
// order as doubles, out[node_index * count * ne + obj * ne + e]
int read_lime(const uint8_t* d, size_t n, int want_type, int want_count, const int* want, double* out,
              milc_io_info* info) {
  std::vector<Rec> recs;
  int rc = lime_scan(d, n, recs);
  if (rc) return rc;
  LimeField f;
  if ((rc = lime_describe(d, recs, f))) return rc;
  if (want_type && f.datatype != want_type) return -2;
  if (want_count > 0 && f.count != want_count) return -2;
  if (want)
    for (int k = 0; k < 4; k++)
      if (f.dims[k] != want[k]) return -4;
  const Lattice L(f.dims);
  const int ne = f.datatype == 1 ? 18 : 6;
  const int es = f.precision == 1 ? 4 : 8;
  const size_t site_bytes = (size_t)f.count * ne * es;
  if (f.bin->len != L.V * site_bytes) return -2;
  milc_io_info inf{};
  inf.format = f.format;
  inf.precision = f.precision;
  for (int k = 0; k < 4; k++) inf.dims[k] = f.dims[k];
  inf.count = f.count;
  inf.datatype = f.datatype;
  if (f.cks) {
    inf.has_checksum = 1;
    inf.stored_a = (unsigned)strtoul(xml_tag(d, f.cks, "suma").c_str(), nullptr, 16);
    inf.stored_b = (unsigned)strtoul(xml_tag(d, f.cks, "sumb").c_str(), nullptr, 16);
  }
  const uint8_t* data = d + f.bin->off;
  const size_t per_site = (size_t)f.count * ne;
  std::vector<uint32_t> pa(16, 0), pb(16, 0);
  parallel_for(L.V, [&](size_t b, size_t e, unsigned tid) {
    uint32_t a = 0, c = 0;
    const Crc32& crc = crc32();
    for (size_t s = b; s < e; s++) {
      const uint8_t* src = data + s * site_bytes;
      uint32_t w = crc(src, site_bytes);
      a ^= rotl(w, (unsigned)(s % 29));
      c ^= rotl(w, (unsigned)(s % 31));
      if (!out) continue;
      double* dst = out + L.node_index(s) * per_site;
      if (es == 4) {
        for (size_t j = 0; j < per_site; j++) {
          uint32_t u = ld32(src + 4 * j, true);
          float v;
          memcpy(&v, &u, 4);
          dst[j] = (double)v;
        }
      } else {
        for (size_t j = 0; j < per_site; j++) {
          uint64_t u = ld64be(src + 8 * j);
          memcpy(&dst[j], &u, 8);
        }
      }
    }
    pa[tid] = a;
    pb[tid] = c;
  });
  for (int t = 0; t < 16; t++) {
    inf.computed_a ^= pa[t];
    inf.computed_b ^= pb[t];
  }
  inf.checksum_ok = inf.has_checksum && inf.computed_a == inf.stored_a && inf.computed_b == inf.stored_b;
  set_info(info, inf);
  if (inf.has_checksum && !inf.checksum_ok) return -3;
  return 0;
}

void put_be16(uint8_t* p, uint16_t v) {
  p[0] = (uint8_t)(v >> 8);
  p[1] = (uint8_t)v;
}

bool lime_put(FILE* fp, const char* type, const void* data, size_t len, bool mb, bool me) {
  uint8_t h[144];
  memset(h, 0, sizeof h);
  uint32_t m = __builtin_bswap32(kLimeMagic);
  memcpy(h, &m, 4);
  put_be16(h + 4, 1);
  put_be16(h + 6, (uint16_t)((mb ? 0x8000 : 0) | (me ? 0x4000 : 0)));
  uint64_t l = __builtin_bswap64((uint64_t)len);
  memcpy(h + 8, &l, 8);
  strncpy((char*)h + 16, type, 127);
  static const uint8_t zeros[8] = {0};
  size_t pad = (8 - len % 8) % 8;
  return fwrite(h, 1, 144, fp) == 144 && (len == 0 || fwrite(data, 1, len, fp) == len) &&
         (pad == 0 || fwrite(zeros, 1, pad, fp) == pad);
}

bool lime_put_str(FILE* fp, const char* type, const std::string& s, bool mb, bool me) {
  return lime_put(fp, type, s.c_str(), s.size() + 1, mb, me);  // with its NUL, as QIO writes
}

std::string date_utc() {
  time_t now = time(nullptr);
  struct tm tmv;
  gmtime_r(&now, &tmv);
  char b[64];
  strftime(b, sizeof b, "%a %b %e %H:%M:%S %Y UTC", &tmv);
  return b;
}

// format: MILC_IO_SCIDAC or MILC_IO_ILDG (gauge only); datatype 1 matrix, 2 vector
int write_lime(const char* path, int format, int precision, int datatype, int count, const int* dims,
               const double* in, milc_io_info* info) {
  const Lattice L(dims);
  const int ne = datatype == 1 ? 18 : 6;
  const int es = precision == 1 ? 4 : 8;
  const size_t per_site = (size_t)count * ne, site_bytes = per_site * es;
  std::unique_ptr<uint8_t[]> bin(new (std::nothrow) uint8_t[L.V * site_bytes]);
  if (!bin) return -1;
  std::vector<uint32_t> pa(16, 0), pb(16, 0);
  parallel_for(L.V, [&](size_t b, size_t e, unsigned tid) {
    uint32_t a = 0, c = 0;
    const Crc32& crc = crc32();
    for (size_t s = b; s < e; s++) {
      const double* src = in + L.node_index(s) * per_site;
      uint8_t* dst = bin.get() + s * site_bytes;
      for (size_t j = 0; j < per_site; j++) {
        if (es == 4) {
          float v = (float)src[j];
          uint32_t u;
          memcpy(&u, &v, 4);
          u = __builtin_bswap32(u);
          memcpy(dst + 4 * j, &u, 4);
        } else {
          uint64_t u;
          memcpy(&u, &src[j], 8);
          u = __builtin_bswap64(u);
          memcpy(dst + 8 * j, &u, 8);
        }
      }
      uint32_t w = crc(dst, site_bytes);
      a ^= rotl(w, (unsigned)(s % 29));
      c ^= rotl(w, (unsigned)(s % 31));
    }
    pa[tid] = a;
    pb[tid] = c;
  });
  uint32_t suma = 0, sumb = 0;
  for (int t = 0; t < 16; t++) {
    suma ^= pa[t];
    sumb ^= pb[t];
  }
  char tmp[512];
  const char* xml = "<?xml version=\"1.0\" encoding=\"UTF-8\"?>";
  snprintf(tmp, sizeof tmp,
           "%s<scidacFile><version>1.1</version><spacetime>4</spacetime><dims>%d %d %d %d </dims>"
           "<volfmt>0</volfmt></scidacFile>",
           xml, dims[0], dims[1], dims[2], dims[3]);
  std::string priv_file = tmp;
  const char* dname = datatype == 1 ? "ColorMatrix" : "ColorVector";
  snprintf(tmp, sizeof tmp,
           "%s<scidacRecord><version>1.0</version><date>%s</date><globaldata>0</globaldata>"
           "<datatype>QDP_%c3_%s</datatype><precision>%c</precision><colors>3</colors>"
           "<typesize>%d</typesize><datacount>%d</datacount></scidacRecord>",
           xml, date_utc().c_str(), precision == 1 ? 'F' : 'D', dname, precision == 1 ? 'F' : 'D', ne * es,
           count);
  std::string priv_rec = tmp;
  snprintf(tmp, sizeof tmp,
           "%s<scidacChecksum><version>1.0</version><suma>%x</suma><sumb>%x</sumb></scidacChecksum>", xml,
           suma, sumb);
  std::string cks = tmp;
  FILE* fp = fopen(path, "wb");
  if (!fp) return -1;
  bool ok = lime_put_str(fp, "scidac-private-file-xml", priv_file, true, false) &&
            lime_put_str(fp, "scidac-file-xml", "milc_qcd_amd lattice field", false, true) &&
            lime_put_str(fp, "scidac-private-record-xml", priv_rec, true, false) &&
            lime_put_str(fp, "scidac-record-xml", datatype == 1 && count == 4 ? "gauge configuration" : "field",
                         false, false);
  if (ok && format == MILC_IO_ILDG) {
    snprintf(tmp, sizeof tmp,
             "%s<ildgFormat xmlns=\"http://www.lqcd.org/ildg\" "
             "xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\" "
             "xsi:schemaLocation=\"http://www.lqcd.org/ildg/filefmt.xsd\"><version>1.0</version>"
             "<field>su3gauge</field><precision>%d</precision><lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt>"
             "</ildgFormat>",
             xml, precision == 1 ? 32 : 64, dims[0], dims[1], dims[2], dims[3]);
    ok = lime_put_str(fp, "ildg-format", tmp, false, false) &&
         lime_put_str(fp, "ildg-data-lfn", path, false, false) &&
         lime_put(fp, "ildg-binary-data", bin.get(), L.V * site_bytes, false, false);
  } else if (ok) {
    ok = lime_put(fp, "scidac-binary-data", bin.get(), L.V * site_bytes, false, false);
  }
  ok = ok && lime_put_str(fp, "scidac-checksum", cks, false, true);
  ok = (fclose(fp) == 0) && ok;
  if (!ok) return -1;
  milc_io_info inf{};
  inf.format = format;
  inf.precision = precision;
  for (int k = 0; k < 4; k++) inf.dims[k] = dims[k];
  inf.count = count;
  inf.datatype = datatype;
  inf.has_checksum = 1;
  inf.checksum_ok = 1;
  inf.stored_a = inf.computed_a = suma;
  inf.stored_b = inf.computed_b = sumb;
  set_info(info, inf);
  return 0;
}

int read_any(const char* path, int want_type, int want_count, const int* want, double* out, milc_io_info* info) {
  if (!path) return -5;
  if (want && !dims_ok(want)) return -5;
  FileBuf b;
  if (read_file(path, b)) return -1;
  if (b.n < 4) return -1;
  uint32_t m;
  memcpy(&m, b.p.get(), 4);
  if (m == kV5Magic || __builtin_bswap32(m) == kV5Magic) {
    if (want_type == 2 || (want_count > 0 && want_count != 4)) return -2;
    return read_v5(b.p.get(), b.n, want, out, info);
  }
  if (__builtin_bswap32(m) == kLimeMagic) return read_lime(b.p.get(), b.n, want_type, want_count, want, out, info);
  return -2;
}

}  // namespace

extern "C" {

int milc_io_read_info(const char* path, milc_io_info* info) {
  if (!info) return -5;
  int rc = read_any(path, 0, 0, nullptr, nullptr, info);
  return rc == -3 ? 0 : rc;  // a checksum mismatch is reported in info
}

int milc_io_read_gauge(const char* path, const int dims[4], double* links, milc_io_info* info) {
  if (!links || !dims) return -5;
  return read_any(path, 1, 4, dims, links, info);
}

int milc_io_write_gauge(const char* path, int format, int precision, const int dims[4], const double* links,
                        milc_io_info* info) {
  if (!path || !links || !dims_ok(dims)) return -5;
  if (format == MILC_IO_V5) return write_v5(path, dims, links, info);
  if ((format != MILC_IO_SCIDAC && format != MILC_IO_ILDG) || (precision != 1 && precision != 2)) return -5;
  return write_lime(path, format, precision, 1, 4, dims, links, info);
}

int milc_io_read_ksvector(const char* path, const int dims[4], double* vec, milc_io_info* info) {
  if (!vec || !dims) return -5;
  return read_any(path, 2, 1, dims, vec, info);
}

int milc_io_write_ksvector(const char* path, int precision, const int dims[4], const double* vec,
                           milc_io_info* info) {
  if (!path || !vec || !dims_ok(dims) || (precision != 1 && precision != 2)) return -5;
  return write_lime(path, MILC_IO_SCIDAC, precision, 2, 1, dims, vec, info);
}

int milc_io_read_colormatrix(const char* path, const int dims[4], int count, double* mat, milc_io_info* info) {
  if (!mat || !dims || count <= 0 || count > 64) return -5;
  FileBuf b;
  if (!path) return -5;
  if (read_file(path, b)) return -1;
  if (b.n < 4) return -1;
  uint32_t m;
  memcpy(&m, b.p.get(), 4);
  if (__builtin_bswap32(m) != kLimeMagic) return -2;  // colour-matrix fields are SciDAC files
  if (!dims_ok(dims)) return -5;
  return read_lime(b.p.get(), b.n, 1, count, dims, mat, info);
}

int milc_io_write_colormatrix(const char* path, int precision, const int dims[4], int count, const double* mat,
                              milc_io_info* info) {
  if (!path || !mat || !dims_ok(dims) || count <= 0 || count > 64 || (precision != 1 && precision != 2)) return -5;
  return write_lime(path, MILC_IO_SCIDAC, precision, 1, count, dims, mat, info);
}

}  // extern "C"
