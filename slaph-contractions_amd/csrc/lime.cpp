// Lime / ILDG gauge-configuration reader (GaugeField.cpp:762-928, :111-149).
//
// A lime file is a sequence of records, each a 144-byte big-endian header
// (magic 0x456789ab, version, flags, 64-bit payload length, 128-byte type
// string) followed by the payload padded to a multiple of 8 bytes.  The
// binary record of an ILDG su3gauge field holds, for t, z, y, x (x fastest),
// the four links mu = x, y, z, t, each 3 x 3 row-major complex numbers of
// big-endian doubles.  The SciDAC checksum record holds suma / sumb: the xor
// over sites of rank r of the site's CRC-32 rotated left by r mod 29 / 31.
#include "lime.hpp"

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace slaph {

namespace {

struct File {
  FILE *fp = nullptr;
  explicit File(std::string const &p) : fp(std::fopen(p.c_str(), "rb")) {}
  ~File() {
    if (fp) std::fclose(fp);
  }
};

uint64_t be64(const unsigned char *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
  return v;
}

uint32_t be32(const unsigned char *p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

double be_double(const unsigned char *p) {
  const uint64_t v = be64(p);
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}

uint32_t rotl(uint32_t c, int n) { return n ? (c << n) | (c >> (32 - n)) : c; }

// The text between <tag> and </tag> of a small XML record ("" if absent).
std::string xml_field(std::string const &xml, std::string const &tag) {
  const std::string open = "<" + tag + ">", close = "</" + tag + ">";
  const size_t a = xml.find(open);
  if (a == std::string::npos) return "";
  const size_t b = xml.find(close, a + open.size());
  if (b == std::string::npos) return "";
  std::string v = xml.substr(a + open.size(), b - a - open.size());
  const size_t s = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
  return s == std::string::npos ? "" : v.substr(s, e - s + 1);
}

struct Record {
  std::string type;
  long long offset;  // payload start
  uint64_t length;
};

}  // namespace

void read_ildg_gauge(std::string const &path, int Lx, int Ly, int Lz, int Lt, int slice_i, int slice_f,
                     std::function<void(int, const std::complex<double> *)> const &sink) {
  if (Lx <= 0 || Ly <= 0 || Lz <= 0 || Lt <= 0) throw std::runtime_error("gauge reader: empty lattice");
  if (slice_i < 0 || slice_f >= Lt || slice_i > slice_f)
    throw std::out_of_range("gauge reader: timeslice range [" + std::to_string(slice_i) + ", " +
                             std::to_string(slice_f) + "] outside [0, " + std::to_string(Lt - 1) + "]");
  File f(path);
  if (!f.fp) throw std::runtime_error("Could not open file " + path);
  // scan the record headers
  std::vector<Record> recs;
  long long pos = 0;
  for (;;) {
    unsigned char h[144];
    if (fseeko(f.fp, pos, SEEK_SET) != 0) break;
    const size_t got = std::fread(h, 1, sizeof h, f.fp);
    if (got == 0) break;
    if (got != sizeof h) throw std::runtime_error("truncated lime record header in file " + path);
    if (be32(h) != 0x456789abu) throw std::runtime_error("not a lime file (bad record magic) " + path);
    Record r;
    r.length = be64(h + 8);
    r.type.assign(reinterpret_cast<const char *>(h + 16), strnlen(reinterpret_cast<const char *>(h + 16), 128));
    r.offset = pos + 144;
    recs.push_back(r);
    pos = r.offset + (long long)((r.length + 7) / 8 * 8);
  }
  auto payload = [&](Record const &r) {
    if (r.length > (1u << 20)) throw std::runtime_error("oversized lime metadata record " + r.type + " in " + path);
    std::string s(r.length, '\0');
    if (fseeko(f.fp, r.offset, SEEK_SET) != 0 || std::fread(&s[0], 1, r.length, f.fp) != r.length)
      throw std::runtime_error("short read of lime record " + r.type + " in " + path);
    return s;
  };
  int bin = -1;
  for (size_t i = 0; i < recs.size() && bin < 0; ++i)
    if (recs[i].type == "ildg-binary-data") bin = int(i);
  if (bin < 0) throw std::runtime_error("no ildg-binary-data record found in file " + path);
  // the format record (before the binary data in ILDG files) and the
  // checksum record (after it)
  std::string suma_s, sumb_s;
  for (size_t i = 0; i < recs.size(); ++i) {
    if (recs[i].type == "ildg-format" && int(i) < bin) {
      const std::string x = payload(recs[i]);
      const std::string dims[4] = {xml_field(x, "lx"), xml_field(x, "ly"), xml_field(x, "lz"), xml_field(x, "lt")};
      const int want_dims[4] = {Lx, Ly, Lz, Lt};
      for (int d = 0; d < 4; ++d)
        if (!dims[d].empty() && std::stol(dims[d]) != want_dims[d])
          throw std::runtime_error("lattice of " + path + " (" + dims[0] + " " + dims[1] + " " + dims[2] + " " +
                                   dims[3] + ") differs from Lx Ly Lz Lt of the input file");
      const std::string prec = xml_field(x, "precision");
      if (!prec.empty() && prec != "64") throw std::runtime_error("Single precision read!");
    }
    if (recs[i].type == "scidac-checksum" && int(i) > bin && suma_s.empty()) {
      const std::string x = payload(recs[i]);
      suma_s = xml_field(x, "suma");
      sumb_s = xml_field(x, "sumb");
    }
  }
  const long long V3 = (long long)Lx * Ly * Lz;
  const uint64_t want = (uint64_t)V3 * Lt * 72 * sizeof(double);
  const Record &b = recs[bin];
  if (b.length != want) {
    if (b.length == want / 2) throw std::runtime_error("Single precision read!");
    throw std::runtime_error("Probably wrong lattice size or precision (bytes=" + std::to_string(b.length) +
                             ") in file " + path + " expected " + std::to_string(want));
  }
  // the binary data, one timeslice at a time
  const size_t site_bytes = 72 * sizeof(double);
  std::vector<unsigned char> raw(size_t(V3) * site_bytes);
  std::vector<std::complex<double>> links(size_t(V3) * 27);
  uint32_t suma = 0, sumb = 0;
  if (fseeko(f.fp, b.offset, SEEK_SET) != 0) throw std::runtime_error("seek failed in " + path);
  for (int t = 0; t < Lt; ++t) {
    if (std::fread(raw.data(), 1, raw.size(), f.fp) != raw.size())
      throw std::runtime_error("short read of the ildg-binary-data record of " + path);
    for (long long s = 0; s < V3; ++s) {
      const uint64_t rank = (uint64_t)t * V3 + s;
      const uint32_t c = uint32_t(crc32(0L, raw.data() + s * site_bytes, uInt(site_bytes)));
      suma ^= rotl(c, int(rank % 29));
      sumb ^= rotl(c, int(rank % 31));
    }
    if (t < slice_i || t > slice_f) continue;
    for (int z = 0; z < Lz; ++z)
      for (int y = 0; y < Ly; ++y)
        for (int x = 0; x < Lx; ++x) {
          const unsigned char *site = raw.data() + (((size_t)z * Ly + y) * Lx + x) * site_bytes;
          const size_t dst = ((size_t)x * Ly + y) * Lz + z;
          for (int mu = 0; mu < 3; ++mu)
            for (int r = 0; r < 3; ++r)
              for (int c = 0; c < 3; ++c) {
                const unsigned char *e = site + ((mu * 9 + r * 3 + c) * 2) * sizeof(double);
                links[(dst * 3 + mu) * 9 + c * 3 + r] = {be_double(e), be_double(e + 8)};
              }
        }
    sink(t, links.data());
  }
  if (!suma_s.empty() || !sumb_s.empty()) {
    const unsigned long ra = std::stoul(suma_s.empty() ? "0" : suma_s, nullptr, 16);
    const unsigned long rb = std::stoul(sumb_s.empty() ? "0" : sumb_s, nullptr, 16);
    if (ra != suma || rb != sumb) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "SciDAC checksum mismatch: computed %08x %08x, record %08lx %08lx in file ", suma,
                    sumb, ra, rb);
      throw std::runtime_error(msg + path);
    }
  }
}

}  // namespace slaph
