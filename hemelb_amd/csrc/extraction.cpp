// Extraction (.xtr v5) and offset (.off v1) files for one rank: property
// output and the binary64 distribution checkpoint.  Host code; the caller
// copies the fields off the device (hlb_get_distributions,
// hlb_get_macroscopic, hlb_get_wall_shear_stress).
#include "hemelb_io.h"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

// io::formats (formats.h:22, extraction.h:12-33, offset.h:19-45)
constexpr uint32_t kHemeLbMagic = 0x686c6221u;
constexpr uint32_t kXtrMagic = 0x78747204u;
constexpr uint32_t kXtrVersion = 5;
constexpr uint32_t kOffMagic = 0x6f666604u;
constexpr uint32_t kOffVersion = 1;
constexpr size_t kMainHeaderLength = 60;

size_t type_size(int t) {
  return (t == HLBX_FLOAT || t == HLBX_INT32 || t == HLBX_UINT32) ? 4 : 8;
}

struct Xdr {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v) {
    u32((uint32_t)(v >> 32));
    u32((uint32_t)v);
  }
  void f32(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u32(u);
  }
  void f64(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    u64(u);
  }
  // XDR string: length, bytes, zero padding to a multiple of four
  // (GetStoredLengthOfString, extraction.h:50-60).
  void str(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
  }
  // A value in the field's type.
  void value(int type, double v) {
    switch (type) {
      case HLBX_FLOAT: f32((float)v); break;
      case HLBX_DOUBLE: f64(v); break;
      case HLBX_INT32: u32((uint32_t)(int32_t)v); break;
      case HLBX_UINT32: u32((uint32_t)v); break;
      case HLBX_INT64: u64((uint64_t)(int64_t)v); break;
      default: u64((uint64_t)v); break;
    }
  }
};

struct XdrIn {
  const unsigned char* p;
  size_t n, i = 0;
  bool ok = true;
  uint32_t u32() {
    if (i + 4 > n) {
      ok = false;
      return 0;
    }
    uint32_t v = ((uint32_t)p[i] << 24) | ((uint32_t)p[i + 1] << 16) |
                 ((uint32_t)p[i + 2] << 8) | (uint32_t)p[i + 3];
    i += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  double f64() {
    const uint64_t u = u64();
    double v;
    std::memcpy(&v, &u, 8);
    return v;
  }
  std::string str() {
    const uint32_t len = u32();
    if (!ok || i + len > n) {
      ok = false;
      return {};
    }
    std::string s((const char*)p + i, len);
    i += (len + 3) / 4 * 4;
    return s;
  }
};

std::string offset_path(const std::string& xtr) {
  // io::formats::offset::ExtractionToOffset (offset.h:51-56)
  const size_t dot = xtr.rfind('.');
  if (dot == std::string::npos) return {};
  return xtr.substr(0, dot) + ".off";
}

bool write_all(FILE* f, const std::vector<unsigned char>& b) {
  return std::fwrite(b.data(), 1, b.size(), f) == b.size();
}

struct Field {
  std::string name;
  int type, n, noff;
  std::vector<double> off;
};

}  // namespace

struct hlbx_writer {
  FILE* f = nullptr;
  int64_t n = 0;
  std::vector<int32_t> coords;
  std::vector<Field> fields;
};

extern "C" {

const char* hlbx_last_error(void) { return g_err.c_str(); }

int hlbx_create(const char* path, double voxel_size_m, const double origin_m[3],
                int64_t n_sites, const int32_t* coords, int32_t n_fields,
                const hlbx_field* fields, hlbx_writer** out) {
  if (!path || !origin_m || !out || (n_sites > 0 && !coords) || n_fields < 0 ||
      (n_fields > 0 && !fields) || n_sites < 0)
    return fail("null or negative argument");
  *out = nullptr;
  const std::string off = offset_path(path);
  if (off.empty()) return fail("cannot split the extension from the extraction file name");
  auto w = new hlbx_writer;
  w->n = n_sites;
  w->coords.assign(coords, coords + 3 * n_sites);
  for (int32_t k = 0; k < n_fields; ++k) {
    const hlbx_field& s = fields[k];
    if (!s.name) {
      delete w;
      return fail("field without a name");
    }
    if (s.type < HLBX_FLOAT || s.type > HLBX_UINT64 || s.n_elements < 1 ||
        !(s.n_offsets == 0 || s.n_offsets == 1 || s.n_offsets == s.n_elements) ||
        (s.n_offsets > 0 && !s.offsets)) {
      delete w;
      // LocalPropertyOutput::CalcSiteWriteLen (LocalPropertyOutput.cc:127-141)
      return fail(std::string("invalid field specification: ") + s.name);
    }
    w->fields.push_back(Field{s.name, s.type, s.n_elements, s.n_offsets,
                              std::vector<double>(s.offsets, s.offsets + s.n_offsets)});
  }
  for (int64_t i = 0; i < 3 * n_sites; ++i)
    if (coords[i] < 0) {
      delete w;
      return fail("site coordinates are written as uint32 and must not be negative");
    }
  // Field headers (extraction.h:62-72) and the main header
  // (LocalPropertyOutput.cc:177-212).
  Xdr fh;
  for (const Field& f : w->fields) {
    fh.str(f.name);
    fh.u32((uint32_t)f.n);
    fh.u32((uint32_t)f.type);
    fh.u32((uint32_t)f.noff);
    for (double o : f.off) fh.value(f.type, o);
  }
  Xdr h;
  h.u32(kHemeLbMagic);
  h.u32(kXtrMagic);
  h.u32(kXtrVersion);
  h.f64(voxel_size_m);
  for (int a = 0; a < 3; ++a) h.f64(origin_m[a]);
  h.u64((uint64_t)n_sites);
  h.u32((uint32_t)w->fields.size());
  h.u32((uint32_t)fh.b.size());
  h.b.insert(h.b.end(), fh.b.begin(), fh.b.end());
  uint64_t site_len = 12;
  for (const Field& f : w->fields) site_len += (uint64_t)f.n * type_size(f.type);
  // Offset file (LocalPropertyOutput.cc:367-392): header with the rank
  // count, then the start of each rank's data and the end of the last.
  Xdr o;
  o.u32(kHemeLbMagic);
  o.u32(kOffMagic);
  o.u32(kOffVersion);
  o.u32(1u);
  o.u64((uint64_t)h.b.size());
  o.u64((uint64_t)h.b.size() + 8u + site_len * (uint64_t)n_sites);
  FILE* fo = std::fopen(off.c_str(), "wbx");
  if (!fo) {
    delete w;
    return fail("cannot create " + off + ": " + std::strerror(errno));
  }
  const bool ok_off = write_all(fo, o.b);
  if (std::fclose(fo) != 0 || !ok_off) {
    delete w;
    return fail("cannot write " + off);
  }
  w->f = std::fopen(path, "wbx");
  if (!w->f) {
    const std::string e = std::strerror(errno);
    delete w;
    return fail(std::string("cannot create ") + path + ": " + e);
  }
  if (!write_all(w->f, h.b)) {
    hlbx_close(w);
    return fail(std::string("cannot write ") + path);
  }
  *out = w;
  return 0;
}

int hlbx_append(hlbx_writer* w, uint64_t time_step, const double* const* data) {
  if (!w || !w->f || (!w->fields.empty() && !data)) return fail("null argument");
  for (size_t k = 0; k < w->fields.size(); ++k)
    if (!data[k] && w->n > 0) return fail("missing data for field " + w->fields[k].name);
  // LocalPropertyOutput::Write (LocalPropertyOutput.cc:245-348)
  Xdr b;
  b.u64(time_step);
  for (int64_t i = 0; i < w->n; ++i) {
    for (int a = 0; a < 3; ++a) b.u32((uint32_t)w->coords[3 * i + a]);
    for (size_t k = 0; k < w->fields.size(); ++k) {
      const Field& f = w->fields[k];
      for (int e = 0; e < f.n; ++e) {
        double v = data[k][i * f.n + e];
        if (f.noff == 1) v -= f.off[0];
        else if (f.noff == f.n) v -= f.off[e];
        b.value(f.type, v);
      }
    }
  }
  if (!write_all(w->f, b.b) || std::fflush(w->f) != 0) return fail("write failed");
  return 0;
}

int hlbx_close(hlbx_writer* w) {
  if (!w) return 0;
  int rc = 0;
  if (w->f && std::fclose(w->f) != 0) rc = fail("close failed");
  delete w;
  return rc;
}

int hlbx_read_distributions(const char* path, int64_t n_sites, const int32_t* coords,
                            int32_t q, int64_t target_time, double* f_out,
                            uint64_t* time_out) {
  if (!path || (n_sites > 0 && (!coords || !f_out)) || q < 1 || n_sites < 0)
    return fail("null or invalid argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::vector<unsigned char> all;
  {
    unsigned char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) all.insert(all.end(), buf, buf + r);
    std::fclose(f);
  }
  XdrIn in{all.data(), all.size()};
  // ReadExtractionHeaders (LocalDistributionInput.cc:172-240)
  const uint32_t m0 = in.u32(), m1 = in.u32(), ver = in.u32();
  if (!in.ok) return fail("truncated extraction header");
  if (m0 != kHemeLbMagic) return fail("file does not start with the HemeLB magic number");
  if (m1 != kXtrMagic) return fail("file does not have the extraction magic number");
  if (ver != kXtrVersion) return fail("unsupported extraction file version");
  (void)in.f64();
  for (int a = 0; a < 3; ++a) (void)in.f64();
  const uint64_t n_file = in.u64();
  const uint32_t nfields = in.u32();
  const uint32_t fh_len = in.u32();
  if (!in.ok || in.i != kMainHeaderLength) return fail("truncated extraction header");
  if (nfields != 1) return fail("a checkpoint holds exactly one field");
  const std::string name = in.str();
  const uint32_t nelem = in.u32(), type = in.u32(), noff = in.u32();
  if (!in.ok || in.i != kMainHeaderLength + fh_len)
    return fail("malformed field header");
  if (type != HLBX_DOUBLE || noff != 0 || nelem != (uint32_t)q)
    return fail("checkpoint field '" + name + "' is not " + std::to_string(q) +
                " binary64 values without offsets");
  if (n_file != (uint64_t)n_sites)
    return fail("checkpoint holds " + std::to_string(n_file) + " sites, expected " +
                std::to_string(n_sites));
  const uint64_t rec_len = 8u + (12u + 8u * (uint64_t)q) * n_file;
  const uint64_t body = all.size() - in.i;
  if (body % rec_len) return fail("file length not consistent with a whole number of records");
  const uint64_t nrec = body / rec_len;
  if (nrec == 0) return fail("checkpoint holds no record");
  auto rec_time = [&](uint64_t k) {
    XdrIn r{all.data() + in.i + k * rec_len, 8};
    return r.u64();
  };
  uint64_t k;
  if (target_time < 0) {
    k = nrec - 1;
  } else {
    // records are in increasing time order: binary search as the reference
    uint64_t lo = 0, len = nrec;
    while (len) {
      const uint64_t half = len / 2, m = lo + half;
      if (rec_time(m) < (uint64_t)target_time) {
        lo = m + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    k = lo;
    if (k >= nrec || rec_time(k) != (uint64_t)target_time)
      return fail("time step " + std::to_string(target_time) + " not found in the checkpoint");
  }
  std::unordered_map<uint64_t, int64_t> where;
  where.reserve((size_t)n_sites * 2);
  auto key = [](uint32_t x, uint32_t y, uint32_t z) {
    return ((uint64_t)(x & 0x1fffff) << 42) | ((uint64_t)(y & 0x1fffff) << 21) |
           (uint64_t)(z & 0x1fffff);
  };
  for (int64_t i = 0; i < n_sites; ++i) {
    if (coords[3 * i] < 0 || coords[3 * i + 1] < 0 || coords[3 * i + 2] < 0 ||
        coords[3 * i] >= (1 << 21) || coords[3 * i + 1] >= (1 << 21) ||
        coords[3 * i + 2] >= (1 << 21))
      return fail("site coordinates out of the range of the file format");
    if (!where.emplace(key(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]), i).second)
      return fail("two sites share the same coordinates");
  }
  XdrIn r{all.data() + in.i + k * rec_len, rec_len};
  const uint64_t t = r.u64();
  std::vector<char> seen(n_sites, 0);
  for (uint64_t s = 0; s < n_file; ++s) {
    const uint32_t x = r.u32(), y = r.u32(), z = r.u32();
    auto it = where.find(key(x, y, z));
    if (it == where.end())
      return fail("checkpoint site (" + std::to_string(x) + "," + std::to_string(y) + "," +
                  std::to_string(z) + ") is not a site of this domain");
    if (seen[it->second]++) return fail("checkpoint holds a site twice");
    double* dst = f_out + it->second * q;
    for (int32_t d = 0; d < q; ++d) dst[d] = r.f64();
  }
  if (!r.ok) return fail("truncated record");
  if (time_out) *time_out = t;
  return 0;
}

}  // extern "C"
