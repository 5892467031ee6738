// cs_dump.cpp -- checkpoint / restart of a cubed-sphere state.
//
// The reference restarts from two files: riemann.in, a text file with time,
// dtcur, n_cc_components and the dump number (read DRV:1527-1547, checked
// against the dump DRV:1618-1647, written DRV:1798-1802), and a Silo dump of
// U, the entropy and the face B of every patch (READDATA DRV:1564-1605,
// WRITEDATA DRV:1845-1930; both routines are absent from the reference tree).
// Here both are one binary file that the host arrays of cs_upload_state /
// cs_download_state go into unchanged:
//
//   bytes 0..4095  header (little endian)
//     0   char[8]  magic "CSDUMP\0\0"
//     8   u32      format version (1)
//     12  u32      header bytes (4096)
//     16  u32      byte-order mark 0x01020304
//     20  u32      sections (5)
//     24  i32[12]  cs_config: ntp nr nghost recon riemann emf precision
//                  rk_order radial_bc radial_mesh imex bc_problem
//     72  f64[10]  cs_config: rmin rmax gamma rhofloor prsfloor cfl grav omega(3)
//     152 f64      time
//     160 f64      dtcur
//     168 i32      n_cc_components (8), i32 0
//     176 i64      filenum
//     184 i64      step
//     192 5 x {char[8] name, u64 offset, u64 count, u64 checksum}
//                  sections U, ent, Bxf, Byf, Bzf (count in float64 words)
//     352 u64      checksum of bytes 0..351
//   then each section at its offset (a multiple of 4096), float64 words.
//
// checksum of words w_0 .. w_{n-1} (bytes taken as little-endian u64): the
// sum modulo 2^64 of f(w_i + (i + 1) * 0x9e3779b97f4a7c15), f the SplitMix64
// finaliser (Steele, Lea & Flood 2014).  f is a bijection, so changing any
// one word changes the sum; the position term catches moved words.
#include "cs_dump.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

namespace cs {

void set_error(const char* fmt, ...);
int validate_config(const cs_config* c);

namespace {

const char kMagic[8] = {'C', 'S', 'D', 'U', 'M', 'P', 0, 0};
const char* const kNames[kDumpSections] = {"U", "ent", "Bxf", "Byf", "Bzf"};
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;
constexpr size_t kHeaderSummed = 352;

inline uint64_t mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
// the running checksum over words [first, first + n) of a section
uint64_t checksum_add(uint64_t sum, const void* words, uint64_t n, uint64_t first) {
  const unsigned char* p = (const unsigned char*)words;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t w;
    memcpy(&w, p + 8 * i, 8);
    sum += mix64(w + (first + i + 1) * kGolden);
  }
  return sum;
}

void put32(unsigned char* h, size_t off, uint32_t v) { memcpy(h + off, &v, 4); }
void put64(unsigned char* h, size_t off, uint64_t v) { memcpy(h + off, &v, 8); }
void putd(unsigned char* h, size_t off, double v) { memcpy(h + off, &v, 8); }
uint32_t get32(const unsigned char* h, size_t off) { uint32_t v; memcpy(&v, h + off, 4); return v; }
uint64_t get64(const unsigned char* h, size_t off) { uint64_t v; memcpy(&v, h + off, 8); return v; }
double getd(const unsigned char* h, size_t off) { double v; memcpy(&v, h + off, 8); return v; }

uint64_t align_up(uint64_t x) { return (x + kDumpHeaderBytes - 1) / kDumpHeaderBytes * kDumpHeaderBytes; }

void encode_config(unsigned char* h, const cs_config* c) {
  const int32_t iv[12] = {c->ntp, c->nr, c->nghost, c->recon, c->riemann, c->emf, c->precision, c->rk_order,
                          c->radial_bc, c->radial_mesh, c->imex, c->bc_problem};
  for (int q = 0; q < 12; ++q) put32(h, 24 + 4 * q, (uint32_t)iv[q]);
  const double dv[10] = {c->rmin, c->rmax, c->gamma, c->rhofloor, c->prsfloor, c->cfl, c->grav,
                         c->omega[0], c->omega[1], c->omega[2]};
  for (int q = 0; q < 10; ++q) putd(h, 72 + 8 * q, dv[q]);
}
void decode_config(const unsigned char* h, cs_config* c) {
  memset(c, 0, sizeof(*c));
  int32_t iv[12];
  for (int q = 0; q < 12; ++q) iv[q] = (int32_t)get32(h, 24 + 4 * q);
  c->ntp = iv[0]; c->nr = iv[1]; c->nghost = iv[2]; c->recon = iv[3]; c->riemann = iv[4]; c->emf = iv[5];
  c->precision = iv[6]; c->rk_order = iv[7]; c->radial_bc = iv[8]; c->radial_mesh = iv[9]; c->imex = iv[10];
  c->bc_problem = iv[11];
  c->rmin = getd(h, 72); c->rmax = getd(h, 80); c->gamma = getd(h, 88); c->rhofloor = getd(h, 96);
  c->prsfloor = getd(h, 104); c->cfl = getd(h, 112); c->grav = getd(h, 120);
  for (int q = 0; q < 3; ++q) c->omega[q] = getd(h, 128 + 8 * q);
}

struct File {
  FILE* f = nullptr;
  ~File() { if (f) fclose(f); }
};

// Reads and checks the header; on success hd holds its bytes and meta the
// decoded fields.  size: the file's length, checked against the sections.
int read_header(FILE* f, const char* path, unsigned char* hd, cs_dump_meta* meta) {
  if (fseeko(f, 0, SEEK_END) != 0) { set_error("%s: cannot seek", path); return CS_E_IO; }
  const off_t size = ftello(f);
  if (size < (off_t)kDumpHeaderBytes || fseeko(f, 0, SEEK_SET) != 0 ||
      fread(hd, 1, kDumpHeaderBytes, f) != kDumpHeaderBytes) {
    set_error("%s: shorter than a dump header", path);
    return CS_E_FORMAT;
  }
  if (memcmp(hd, kMagic, 8) != 0) { set_error("%s: not a cubed-sphere dump (magic)", path); return CS_E_FORMAT; }
  if (get32(hd, 16) != 0x01020304u) { set_error("%s: written with the other byte order", path); return CS_E_FORMAT; }
  const uint32_t version = get32(hd, 8);
  if (version != kDumpVersion) {
    set_error("%s: format version %u, this library reads version %u", path, version, kDumpVersion);
    return CS_E_FORMAT;
  }
  if (get32(hd, 12) != kDumpHeaderBytes || get32(hd, 20) != kDumpSections) {
    set_error("%s: header size or section count not those of version %u", path, kDumpVersion);
    return CS_E_FORMAT;
  }
  if (checksum_add(0, hd, kHeaderSummed / 8, 0) != get64(hd, kHeaderSummed)) {
    set_error("%s: header checksum mismatch", path);
    return CS_E_FORMAT;
  }
  memset(meta, 0, sizeof(*meta));
  decode_config(hd, &meta->cfg);
  meta->time = getd(hd, 152);
  meta->dtcur = getd(hd, 160);
  meta->n_cc_components = (int32_t)get32(hd, 168);
  meta->filenum = (int64_t)get64(hd, 176);
  meta->step = (int64_t)get64(hd, 184);
  meta->version = (int32_t)version;
  if (meta->n_cc_components != 8) {
    // the reference stops the same way on a component count it was not built for (DRV:1641-1646)
    set_error("%s: n_cc_components %d, the state has 8", path, meta->n_cc_components);
    return CS_E_FORMAT;
  }
  if (validate_config(&meta->cfg) != CS_OK) {
    set_error("%s: the configuration in the header is not valid", path);
    return CS_E_FORMAT;
  }
  uint64_t counts[kDumpSections];
  dump_counts(&meta->cfg, counts);
  for (int q = 0; q < kDumpSections; ++q) {
    const unsigned char* s = hd + 192 + 32 * q;
    char nm[9] = {0};
    memcpy(nm, s, 8);
    const uint64_t off = get64(s, 8), cnt = get64(s, 16);
    if (strcmp(nm, kNames[q]) != 0 || cnt != counts[q] || off % kDumpHeaderBytes != 0 || off < kDumpHeaderBytes ||
        off + 8 * cnt > (uint64_t)size) {
      set_error("%s: section %d (%s) does not fit the header's grid or the file's length", path, q, kNames[q]);
      return CS_E_FORMAT;
    }
  }
  return CS_OK;
}

}  // namespace

void dump_counts(const cs_config* c, uint64_t counts[kDumpSections]) {
  const uint64_t nc = (uint64_t)c->ntp + 2 * c->nghost, nrc = (uint64_t)c->nr + 2 * c->nghost;
  const uint64_t cells = nc * nc * nrc * 6, faces = (nc + 1) * (nc + 1) * (nrc + 1) * 6;
  counts[0] = 8 * cells;
  counts[1] = cells;
  counts[2] = counts[3] = counts[4] = faces;
}

int dump_write(const char* path, const cs_config* cfg, const cs_dump_meta* meta, DumpIo& io) {
  if (!path || !cfg || !meta) { set_error("cs dump: null argument"); return CS_E_INVALID; }
  int rc = validate_config(cfg);
  if (rc) return rc;
  if (io.chunk_n == 0 || io.chunk_n % 4 != 0) { set_error("cs dump: bad staging size"); return CS_E_INVALID; }
  uint64_t counts[kDumpSections], offs[kDumpSections], sums[kDumpSections];
  dump_counts(cfg, counts);
  uint64_t at = kDumpHeaderBytes;
  for (int q = 0; q < kDumpSections; ++q) {
    offs[q] = at;
    at = align_up(at + 8 * counts[q]);
  }
  const std::string tmp = std::string(path) + ".tmp";
  File F;
  F.f = fopen(tmp.c_str(), "wb");
  if (!F.f) { set_error("%s: cannot create", tmp.c_str()); return CS_E_IO; }
  std::vector<unsigned char> hd(kDumpHeaderBytes, 0);
  std::vector<unsigned char> zeros(kDumpHeaderBytes, 0);
  if (fwrite(hd.data(), 1, hd.size(), F.f) != hd.size()) { set_error("%s: write failed", tmp.c_str()); return CS_E_IO; }
  for (int q = 0; q < kDumpSections; ++q) {
    uint64_t sum = 0;
    for (uint64_t off = 0; off < counts[q]; off += io.chunk_n) {
      const uint64_t n = std::min<uint64_t>(io.chunk_n, counts[q] - off);
      if ((rc = io.get(q, off, n, io.chunk))) return rc;
      sum = checksum_add(sum, io.chunk, n, off);
      if (fwrite(io.chunk, 8, n, F.f) != n) { set_error("%s: write failed", tmp.c_str()); return CS_E_IO; }
    }
    sums[q] = sum;
    const uint64_t end = offs[q] + 8 * counts[q];
    const uint64_t pad = (q + 1 < kDumpSections ? offs[q + 1] : align_up(end)) - end;
    if (pad && fwrite(zeros.data(), 1, pad, F.f) != pad) { set_error("%s: write failed", tmp.c_str()); return CS_E_IO; }
  }
  memcpy(hd.data(), kMagic, 8);
  put32(hd.data(), 8, kDumpVersion);
  put32(hd.data(), 12, (uint32_t)kDumpHeaderBytes);
  put32(hd.data(), 16, 0x01020304u);
  put32(hd.data(), 20, kDumpSections);
  encode_config(hd.data(), cfg);
  putd(hd.data(), 152, meta->time);
  putd(hd.data(), 160, meta->dtcur);
  put32(hd.data(), 168, 8);
  put64(hd.data(), 176, (uint64_t)meta->filenum);
  put64(hd.data(), 184, (uint64_t)meta->step);
  for (int q = 0; q < kDumpSections; ++q) {
    unsigned char* s = hd.data() + 192 + 32 * q;
    memcpy(s, kNames[q], strlen(kNames[q]));
    put64(s, 8, offs[q]);
    put64(s, 16, counts[q]);
    put64(s, 24, sums[q]);
  }
  put64(hd.data(), kHeaderSummed, checksum_add(0, hd.data(), kHeaderSummed / 8, 0));
  if (fseeko(F.f, 0, SEEK_SET) != 0 || fwrite(hd.data(), 1, hd.size(), F.f) != hd.size() || fflush(F.f) != 0 ||
      fsync(fileno(F.f)) != 0) {
    set_error("%s: write failed", tmp.c_str());
    return CS_E_IO;
  }
  fclose(F.f);
  F.f = nullptr;
  if (rename(tmp.c_str(), path) != 0) { set_error("%s: cannot rename into place", path); return CS_E_IO; }
  return CS_OK;
}

int dump_read(const char* path, const cs_config* cfg, cs_dump_meta* meta, DumpIo& io) {
  if (!path || !meta) { set_error("cs dump: null argument"); return CS_E_INVALID; }
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) { set_error("%s: cannot open", path); return CS_E_IO; }
  std::vector<unsigned char> hd(kDumpHeaderBytes);
  int rc = read_header(F.f, path, hd.data(), meta);
  if (rc) return rc;
  if (cfg) {
    const cs_config& a = meta->cfg;
    if (a.ntp != cfg->ntp || a.nr != cfg->nr || a.nghost != cfg->nghost || a.radial_mesh != cfg->radial_mesh ||
        a.rmin != cfg->rmin || a.rmax != cfg->rmax) {
      set_error("%s: grid %dx%d^2x%d (nghost %d, radii %g..%g, mesh %d) is not the caller's %dx%d^2x%d "
                "(nghost %d, radii %g..%g, mesh %d)", path, 6, a.ntp, a.nr, a.nghost, a.rmin, a.rmax,
                a.radial_mesh, 6, cfg->ntp, cfg->nr, cfg->nghost, cfg->rmin, cfg->rmax, cfg->radial_mesh);
      return CS_E_INVALID;
    }
  }
  if (!io.put) return CS_OK;
  if (io.chunk_n == 0 || io.chunk_n % 4 != 0) { set_error("cs dump: bad staging size"); return CS_E_INVALID; }
  for (int q = 0; q < kDumpSections; ++q) {
    const unsigned char* s = hd.data() + 192 + 32 * q;
    const uint64_t off0 = get64(s, 8), cnt = get64(s, 16), want = get64(s, 24);
    if (fseeko(F.f, (off_t)off0, SEEK_SET) != 0) { set_error("%s: cannot seek", path); return CS_E_IO; }
    uint64_t sum = 0;
    for (uint64_t off = 0; off < cnt; off += io.chunk_n) {
      const uint64_t n = std::min<uint64_t>(io.chunk_n, cnt - off);
      if (fread(io.chunk, 8, n, F.f) != n) { set_error("%s: short read in section %s", path, kNames[q]); return CS_E_IO; }
      sum = checksum_add(sum, io.chunk, n, off);
      if ((rc = io.put(q, off, n, io.chunk))) return rc;
    }
    if (sum != want) {
      set_error("%s: checksum mismatch in section %s", path, kNames[q]);
      return CS_E_FORMAT;
    }
  }
  return CS_OK;
}

}  // namespace cs

using namespace cs;

namespace {
constexpr uint64_t kHostChunk = 1u << 20;  // 8 MiB of float64 per transfer
}

extern "C" int cs_dump_write(const char* path, const cs_config* cfg, const cs_dump_meta* meta, const double* U,
                             const double* ent, const double* Bxf, const double* Byf, const double* Bzf) {
  if (!U || !ent || !Bxf || !Byf || !Bzf) { set_error("cs_dump_write: null array"); return CS_E_INVALID; }
  const double* src[kDumpSections] = {U, ent, Bxf, Byf, Bzf};
  std::vector<double> chunk(kHostChunk);
  DumpIo io;
  io.chunk = chunk.data();
  io.chunk_n = kHostChunk;
  io.get = [&](int q, uint64_t off, uint64_t n, double* dst) {
    memcpy(dst, src[q] + off, 8 * n);
    return CS_OK;
  };
  return dump_write(path, cfg, meta, io);
}

extern "C" int cs_dump_read(const char* path, const cs_config* cfg, cs_dump_meta* meta, double* U, double* ent,
                            double* Bxf, double* Byf, double* Bzf) {
  if (!cfg) { set_error("cs_dump_read: null configuration"); return CS_E_INVALID; }
  if (!U || !ent || !Bxf || !Byf || !Bzf) { set_error("cs_dump_read: null array"); return CS_E_INVALID; }
  double* dst[kDumpSections] = {U, ent, Bxf, Byf, Bzf};
  std::vector<double> chunk(kHostChunk);
  DumpIo io;
  io.chunk = chunk.data();
  io.chunk_n = kHostChunk;
  io.put = [&](int q, uint64_t off, uint64_t n, const double* src) {
    memcpy(dst[q] + off, src, 8 * n);
    return CS_OK;
  };
  return dump_read(path, cfg, meta, io);
}

extern "C" int cs_dump_info(const char* path, cs_dump_meta* meta) {
  DumpIo io;
  return dump_read(path, nullptr, meta, io);
}
