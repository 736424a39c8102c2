// ILDG / LIME gauge-configuration reader (read_ildg_gauge_conf, io/reader.hpp:75-98;
// record search in io/ILDG_File.cpp; checksum in io/checksum.hpp).
//
// A LIME file is a sequence of records, each a 144-byte big-endian header (magic
// 0x456789ab u32, version u16, flags u16, data length u64, type string of 128
// bytes) followed by the data, padded to a multiple of 8 bytes.  The gauge field is
// the "ildg-binary-data" record: per site (t slowest, then z, y, x fastest) four
// su3 in direction order x, y, z, t, big-endian reals of the precision that the
// "ildg-format" record names.  The SciDAC checksum XORs, over the file sites i,
// the CRC-32 of the site's bytes rotated left by i % 29 (suma) and i % 31 (sumb).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nq_internal.hpp"

namespace {

using nq::fail;

uint64_t be(const unsigned char* p, int n)
{
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 8) | p[i];
  return v;
}

// CRC-32 (reflected polynomial 0xedb88320, initial value and final XOR 0xffffffff)
uint32_t crc32(const unsigned char* p, size_t n)
{
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int j = 0; j < 8; j++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : (c >> 1);
      table[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xffffffffu;
  for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c ^ 0xffffffffu;
}

uint32_t rotl(uint32_t v, uint32_t r) { return r ? (v << r) | (v >> (32 - r)) : v; }

// integer between <tag> and </tag>, or -1
long long xml_int(const std::string& s, const char* tag, int base = 10)
{
  const std::string open = std::string("<") + tag + ">";
  const size_t a = s.find(open);
  if (a == std::string::npos) return -1;
  return std::strtoll(s.c_str() + a + open.size(), nullptr, base);
}

}  // namespace

// The sites of one binary record into an lx buffer of nreals doubles per site.
// gauge: the record is "ildg-binary-data", its precision and extents come from
// "ildg-format", and the four directions move from x,y,z,t to nissa's t,x,y,z.
// Otherwise the named record holds nreals reals per site in nissa's order, and
// the precision follows from its length.
static int read_sites(const char* path, const char* record, bool gauge, int nreals, int T, int LX, int LY, int LZ,
                      double* out_lx, uint32_t checks[4])
{
  if (!path || !out_lx || T < 1 || LX < 1 || LY < 1 || LZ < 1) return fail(NQ_EINVAL, "LIME reader: bad argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(NQ_EINVAL, "cannot open %s", path);
  f.seekg(0, std::ios::end);
  const uint64_t fsize = (uint64_t)f.tellg();
  f.seekg(0, std::ios::beg);
  // walk the record headers and read only the payloads needed: the data record
  // (one allocation of its exact size), the format and the checksum after it
  std::vector<unsigned char> payload;
  const unsigned char* data = nullptr;
  uint64_t ndata = 0;
  std::string format, checksum;
  uint64_t pos = 0;
  unsigned char hdr[144];
  while (pos + 144 <= fsize) {
    f.seekg((std::streamoff)pos);
    if (!f.read(reinterpret_cast<char*>(hdr), 144)) return fail(NQ_EINVAL, "%s: read error at byte %llu", path, (unsigned long long)pos);
    if (be(hdr, 4) != 0x456789abu) return fail(NQ_EINVAL, "%s: bad LIME magic at byte %llu", path, (unsigned long long)pos);
    const uint64_t len = be(hdr + 8, 8);
    const std::string type(reinterpret_cast<const char*>(hdr + 16), strnlen(reinterpret_cast<const char*>(hdr + 16), 128));
    const uint64_t start = pos + 144;
    if (len > fsize - start) return fail(NQ_EINVAL, "%s: truncated record %s", path, type.c_str());
    const bool want_data = type == record && !data;
    const bool want_fmt = !want_data && gauge && type == "ildg-format";
    const bool want_chk = !want_data && type == "scidac-checksum" && data && checksum.empty();  // the one after the data
    if (want_data || want_fmt || want_chk) {
      std::vector<unsigned char> buf(len);
      if (len && !f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)len))
        return fail(NQ_EINVAL, "%s: read error in record %s", path, type.c_str());
      if (want_data) {
        payload.swap(buf);
        data = payload.data();
        ndata = len;
      } else if (want_fmt) {
        format.assign(buf.begin(), buf.end());
      } else {
        checksum.assign(buf.begin(), buf.end());
      }
    }
    pos = start + ((len + 7) / 8) * 8;
  }
  if (!data) return fail(NQ_EINVAL, "%s: record %s not found", path, record);
  const int64_t V = (int64_t)T * LX * LY * LZ;
  int prec = (int)xml_int(format, "precision");
  if (prec < 0) prec = ndata == (uint64_t)V * nreals * 4 ? 32 : 64;
  if (prec != 32 && prec != 64) return fail(NQ_EINVAL, "%s: precision %d not supported", path, prec);
  const long long fl[4] = {xml_int(format, "lt"), xml_int(format, "lx"), xml_int(format, "ly"), xml_int(format, "lz")};
  const int want[4] = {T, LX, LY, LZ};
  for (int k = 0; k < 4; k++)
    if (fl[k] >= 0 && fl[k] != want[k])
      return fail(NQ_EINVAL, "%s: lattice %lldx%lldx%lldx%lld in the file, %dx%dx%dx%d requested", path, fl[0], fl[1],
                  fl[2], fl[3], T, LX, LY, LZ);
  const int wb = prec / 8;
  const size_t site_bytes = (size_t)nreals * wb;
  if (ndata != (uint64_t)V * site_bytes)
    return fail(NQ_EINVAL, "%s: %llu data bytes, %lld expected", path, (unsigned long long)ndata, (long long)(V * site_bytes));
  const int per_block = gauge ? 18 : nreals;
  uint32_t suma = 0, sumb = 0;
  for (int64_t i = 0; i < V; i++) {
    const unsigned char* sp = data + i * site_bytes;
    const int64_t x = i % LX, y = (i / LX) % LY, z = (i / ((int64_t)LX * LY)) % LZ, t = i / ((int64_t)LX * LY * LZ);
    const int64_t lx = ((t * LX + x) * LY + y) * LZ + z;
    for (int b = 0; b < nreals / per_block; b++) {
      // gauge: file direction b (x,y,z,t) -> nissa (b + 1) % 4 (t,x,y,z)
      double* u = out_lx + lx * nreals + (gauge ? (b + 1) % 4 : b) * per_block;
      for (int k = 0; k < per_block; k++) {
        const uint64_t w = be(sp + (b * per_block + k) * wb, wb);
        if (wb == 8) {
          std::memcpy(&u[k], &w, 8);
        } else {
          const uint32_t w32 = (uint32_t)w;
          float v;
          std::memcpy(&v, &w32, 4);
          u[k] = v;
        }
      }
    }
    const uint32_t c = crc32(sp, site_bytes);
    suma ^= rotl(c, (uint32_t)(i % 29));
    sumb ^= rotl(c, (uint32_t)(i % 31));
  }
  if (checks) {
    checks[0] = (uint32_t)(checksum.empty() ? 0 : xml_int(checksum, "suma", 16));
    checks[1] = (uint32_t)(checksum.empty() ? 0 : xml_int(checksum, "sumb", 16));
    checks[2] = suma;
    checks[3] = sumb;
  }
  return NQ_OK;
}

extern "C" int nq_read_ildg_gauge_conf(const char* path, int T, int LX, int LY, int LZ, double* conf_lx,
                                       uint32_t checks[4])
{
  return read_sites(path, "ildg-binary-data", true, 72, T, LX, LY, LZ, conf_lx, checks);
}

extern "C" int nq_read_lime_real_vector(const char* path, const char* record, int nreals_per_site, int T, int LX,
                                        int LY, int LZ, double* out_lx, uint32_t checks[4])
{
  if (!record || nreals_per_site < 1) return fail(NQ_EINVAL, "nq_read_lime_real_vector: bad argument");
  return read_sites(path, record, false, nreals_per_site, T, LX, LY, LZ, out_lx, checks);
}

// unitarity_check_lx_conf (operations/gaugeconf.hpp:204-245) with
// su3_get_non_unitariness (new_types/su3_op.hpp:3037-3055): per link the Frobenius
// norm of U U^+ - 1, i.e. sqrt(sum_i (|row_i|^2 - 1)^2 + 2 sum_{i<j} |<row_i, row_j>|^2);
// out[0] = average over the 4 V links, out[1] = maximum; nbroken counts links above 1e-13
extern "C" int nq_host_unitarity_check_lx(int T, int LX, int LY, int LZ, const double* conf_lx, double out[2],
                                          int64_t* nbroken)
{
  if (!conf_lx || !out || T < 1 || LX < 1 || LY < 1 || LZ < 1) return fail(NQ_EINVAL, "nq_host_unitarity_check_lx: bad argument");
  const int64_t nlinks = (int64_t)T * LX * LY * LZ * 4;
  double sum = 0, mx = 0;
  int64_t nb = 0;
  for (int64_t l = 0; l < nlinks; l++) {
    const double* u = conf_lx + l * 18;
    double res = 0;
    for (int i = 0; i < 3; i++) {
      double n2 = 0;
      for (int k = 0; k < 6; k++) n2 += u[i * 6 + k] * u[i * 6 + k];
      res += (n2 - 1.0) * (n2 - 1.0);
      for (int j = i + 1; j < 3; j++) {
        double re = 0, im = 0;  // sum_c conj(u_ic) u_jc
        for (int c = 0; c < 3; c++) {
          const double ar = u[i * 6 + 2 * c], ai = u[i * 6 + 2 * c + 1];
          const double br = u[j * 6 + 2 * c], bi = u[j * 6 + 2 * c + 1];
          re += ar * br + ai * bi;
          im += ar * bi - ai * br;
        }
        res += 2 * (re * re + im * im);
      }
    }
    const double err = std::sqrt(res);
    sum += err;
    mx = err > mx ? err : mx;
    nb += err > 1e-13;
  }
  out[0] = sum / (double)nlinks;
  out[1] = mx;
  if (nbroken) *nbroken = nb;
  return NQ_OK;
}
