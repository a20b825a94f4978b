// gt_restart.cpp -- the restart fields of GETM's NetCDF hotstart file
// (src/ncdf/create_restart_ncdf.F90, write_restart_ncdf.F90,
// read_restart_ncdf.F90) in the NetCDF classic format, written and read
// natively (no NetCDF library): dimensions xax, yax, zax (= kmax+1), the
// scalars of create_restart_ncdf.F90:94-104 (loop, julianday, secondsofday
// as int, timestep as double), one FP64 variable per field over (yax, xax)
// or (zax, yax, xax) in the file's C order (the Fortran (x, y[, z]) order of
// the reference), the global attribute title.  Writing: interior cells
// imin..imax, jmin..jmax, or the whole array with the halos (the reference's
// _WRITE_HOT_HALOS_ build).  Reading (read_restart_ncdf.F90:71-118): the four
// scalars are required, as open_restart_ncdf.F90:48-62 requires them; the
// interior cells are read from an interior-sized file or, with halos (the
// __READ_HOT_HALOS_ build), also from a halo-inclusive one, whose halo
// values are skipped: the caller fills halos (update_2d/3d_halo), as the
// reference does.  Version 1 of the format (32-bit offsets) when the file
// fits in 2 GiB, else version 2 (64-bit offsets), as NF90_64BIT_OFFSET would
// write it; both are read.
//
// Host code only: the device entry points (gt_api.hip) copy through host
// buffers.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/getm_transport.h"

// the error text of gt_last_error (gt_api.hip)
int gt_internal_set_error(int code, const char* msg);

namespace {

enum : uint32_t { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum : uint32_t { NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6 };

int fail(int code, const char* fmt, const char* a = "", const char* b = "") {
  char msg[512];
  std::snprintf(msg, sizeof msg, fmt, a, b);
  return gt_internal_set_error(code, msg);
}

// ---- big-endian encoding -----------------------------------------------------
struct Out {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
  }
};

struct In {
  const uint8_t* p;
  size_t n, o = 0;
  bool ok = true;
  uint32_t u32() {
    if (o + 4 > n) {
      ok = false;
      return 0;
    }
    uint32_t v = 0;
    for (int q = 0; q < 4; q++) v = (v << 8) | p[o++];
    return v;
  }
  uint64_t u64() {
    uint64_t hi = u32();
    return (hi << 32) | u32();
  }
  std::string name() {
    const uint32_t len = u32();
    if (!ok || o + len > n) {
      ok = false;
      return {};
    }
    std::string s((const char*)p + o, len);
    o += (len + 3) & ~3u;
    return s;
  }
};

void put_be_doubles(uint8_t* dst, const double* src, size_t n) {
  for (size_t q = 0; q < n; q++) {
    uint64_t v;
    std::memcpy(&v, &src[q], 8);
    for (int s = 0; s < 8; s++) dst[8 * q + s] = (uint8_t)(v >> (56 - 8 * s));
  }
}

void get_be_doubles(double* dst, const uint8_t* src, size_t n) {
  for (size_t q = 0; q < n; q++) {
    uint64_t v = 0;
    for (int s = 0; s < 8; s++) v = (v << 8) | src[8 * q + s];
    std::memcpy(&dst[q], &v, 8);
  }
}

struct Shape {
  long NX, NY, NK;     // array (halo-inclusive) extents
  long x0, y0, xl, yl; // first cell and count along i, j in the file
};

Shape shape(int imin, int imax, int jmin, int jmax, int kmax, int halos) {
  Shape s;
  s.NX = imax - imin + 5;
  s.NY = jmax - jmin + 5;
  s.NK = kmax + 1;
  s.x0 = halos ? 0 : 2;
  s.y0 = halos ? 0 : 2;
  s.xl = halos ? s.NX : s.NX - 4;
  s.yl = halos ? s.NY : s.NY - 4;
  return s;
}

const char* kTitle = "GETM NetCDF hotstart file";

}  // namespace

extern "C" int gt_restart_write_host(const char* path, int imin, int imax, int jmin, int jmax,
                                     int kmax, int with_halos, const gt_restart_time* tm, int n,
                                     const gt_restart_field* f) {
  const gt_restart_time t0 = {0, 0, 0, 0.0};
  const gt_restart_time& T = tm ? *tm : t0;
  if (!path || n < 0 || (n > 0 && !f) || imax < imin || jmax < jmin || kmax < 0)
    return fail(GT_EARG, "gt_restart_write: bad argument%s%s");
  const Shape s = shape(imin, imax, jmin, jmax, kmax, with_halos);
  for (int v = 0; v < n; v++)
    if (!f[v].name || !f[v].data || (f[v].ndim != 2 && f[v].ndim != 3))
      return fail(GT_EARG, "gt_restart_write: field %s: bad name, data or ndim%s",
                  f[v].name ? f[v].name : "?");
  const uint64_t plane = (uint64_t)s.xl * s.yl * 8;
  uint64_t data = 4 + 4 + 4 + 8;  // loop, julianday, secondsofday, timestep
  for (int v = 0; v < n; v++) data += f[v].ndim == 3 ? plane * s.NK : plane;
  // the header; its size depends only on the format version (4- or 8-byte
  // offsets), so it is built once to measure and once with the offsets
  auto build = [&](int version, uint64_t hdr) {
    Out h;
    h.b.insert(h.b.end(), {'C', 'D', 'F', (uint8_t)version});
    h.u32(0);  // numrecs: no record dimension
    h.u32(NC_DIMENSION);
    h.u32(3);
    h.name("xax");
    h.u32((uint32_t)s.xl);
    h.name("yax");
    h.u32((uint32_t)s.yl);
    h.name("zax");
    h.u32((uint32_t)s.NK);
    h.u32(NC_ATTRIBUTE);
    h.u32(1);
    h.name("title");
    h.u32(NC_CHAR);
    h.u32((uint32_t)std::strlen(kTitle));
    h.b.insert(h.b.end(), kTitle, kTitle + std::strlen(kTitle));
    while (h.b.size() % 4) h.b.push_back(0);
    h.u32(NC_VARIABLE);
    h.u32((uint32_t)n + 4);
    uint64_t off = hdr;
    auto var = [&](const std::string& nm, int ndim, uint32_t type, uint64_t bytes) {
      h.name(nm);
      h.u32((uint32_t)ndim);
      // dimension ids in C order: (zax,) yax, xax
      if (ndim == 3) h.u32(2);
      if (ndim >= 2) {
        h.u32(1);
        h.u32(0);
      }
      h.u32(0);
      h.u32(0);  // no variable attributes
      h.u32(type);
      h.u32(bytes > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)bytes);
      if (version == 1)
        h.u32((uint32_t)off);
      else
        h.u64(off);
      off += bytes;
    };
    var("loop", 0, NC_INT, 4);
    var("julianday", 0, NC_INT, 4);
    var("secondsofday", 0, NC_INT, 4);
    var("timestep", 0, NC_DOUBLE, 8);
    for (int v = 0; v < n; v++)
      var(f[v].name, f[v].ndim, NC_DOUBLE, f[v].ndim == 3 ? plane * s.NK : plane);
    return h;
  };
  const int version = build(1, 0).b.size() + data > 0x7FFFFFFFull ? 2 : 1;
  const Out h = build(version, build(version, 0).b.size());
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return fail(GT_EARG, "gt_restart_write: cannot create %s%s", path);
  bool ok = std::fwrite(h.b.data(), 1, h.b.size(), fp) == h.b.size();
  uint8_t sc[20];
  const int32_t ints[3] = {T.loop, T.julianday, T.secondsofday};
  for (int q = 0; q < 3; q++)
    for (int b = 0; b < 4; b++) sc[4 * q + b] = (uint8_t)((uint32_t)ints[q] >> (24 - 8 * b));
  put_be_doubles(sc + 12, &T.timestep, 1);
  ok = ok && std::fwrite(sc, 1, 20, fp) == 20;
  std::vector<uint8_t> row((size_t)s.xl * 8);
  for (int v = 0; v < n && ok; v++) {
    const long nk = f[v].ndim == 3 ? s.NK : 1;
    for (long k = 0; k < nk && ok; k++)
      for (long j = 0; j < s.yl && ok; j++) {
        const double* src = f[v].data + (k * s.NY + (s.y0 + j)) * s.NX + s.x0;
        put_be_doubles(row.data(), src, (size_t)s.xl);
        ok = std::fwrite(row.data(), 1, row.size(), fp) == row.size();
      }
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) return fail(GT_EARG, "gt_restart_write: write to %s failed%s", path);
  return GT_OK;
}

extern "C" int gt_restart_read_host(const char* path, int imin, int imax, int jmin, int jmax,
                                    int kmax, int with_halos, gt_restart_time* tm, int n,
                                    const gt_restart_field* f) {
  if (!path || n < 0 || (n > 0 && !f)) return fail(GT_EARG, "gt_restart_read: bad argument%s%s");
  // the cells filled: always the interior (read_restart_ncdf.F90:83-101)
  const Shape s = shape(imin, imax, jmin, jmax, kmax, 0);
  const Shape sh = shape(imin, imax, jmin, jmax, kmax, 1);
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return fail(GT_EARG, "gt_restart_read: cannot open %s%s", path);
  // the header: read a generous prefix, grow if it does not parse
  std::vector<uint8_t> buf;
  struct Var {
    uint32_t type;
    std::vector<uint32_t> dims;
    uint64_t begin;
  };
  std::map<std::string, Var> vars;
  std::vector<std::pair<std::string, uint64_t>> dims;
  int version = 0;
  for (size_t want = 1 << 16;; want *= 4) {
    buf.resize(want);
    std::fseek(fp, 0, SEEK_SET);
    const size_t got = std::fread(buf.data(), 1, want, fp);
    In in{buf.data(), got};
    vars.clear();
    dims.clear();
    bool eof_limited = got < want;
    if (got < 8 || std::memcmp(buf.data(), "CDF", 3) != 0 || (buf[3] != 1 && buf[3] != 2)) {
      std::fclose(fp);
      return fail(GT_EARG, "gt_restart_read: %s is not a NetCDF classic file%s", path);
    }
    version = buf[3];
    in.o = 4;
    in.u32();  // numrecs
    uint32_t tag = in.u32(), cnt = in.u32();
    if (tag == NC_DIMENSION)
      for (uint32_t q = 0; q < cnt && in.ok; q++) {
        std::string nm = in.name();
        dims.emplace_back(nm, in.u32());
      }
    auto skip_atts = [&]() {
      uint32_t t = in.u32(), c = in.u32();
      if (t != NC_ATTRIBUTE) return;
      for (uint32_t q = 0; q < c && in.ok; q++) {
        in.name();
        const uint32_t type = in.u32(), ne = in.u32();
        const uint32_t sz = type == 1 || type == 2 ? 1 : type == 3 ? 2 : type == 6 ? 8 : 4;
        in.o += ((uint64_t)ne * sz + 3) & ~3ull;
      }
    };
    skip_atts();
    tag = in.u32();
    cnt = in.u32();
    if (tag == NC_VARIABLE)
      for (uint32_t q = 0; q < cnt && in.ok; q++) {
        Var v;
        std::string nm = in.name();
        const uint32_t nd = in.u32();
        for (uint32_t r = 0; r < nd; r++) v.dims.push_back(in.u32());
        skip_atts();
        v.type = in.u32();
        in.u32();  // vsize
        v.begin = version == 1 ? in.u32() : in.u64();
        vars[nm] = v;
      }
    if (in.ok && in.o <= got) break;
    if (eof_limited || want > (1u << 28)) {
      std::fclose(fp);
      return fail(GT_EARG, "gt_restart_read: %s: header does not parse%s", path);
    }
  }
  auto dimlen = [&](const char* nm) -> long {
    for (auto& d : dims)
      if (d.first == nm) return (long)d.second;
    return -1;
  };
  // an interior-sized file, or with halos also a halo-inclusive one whose
  // first HALO rows and columns are skipped
  long fxl = s.xl, fx0 = 0;
  if (dimlen("zax") != s.NK) fxl = -1;
  if (with_halos && dimlen("xax") == sh.xl && dimlen("yax") == sh.yl) fx0 = 2, fxl = sh.xl;
  else if (dimlen("xax") != s.xl || dimlen("yax") != s.yl) fxl = -1;
  if (fxl < 0) {
    std::fclose(fp);
    return fail(GT_EARG, "gt_restart_read: %s: xax/yax/zax do not match the domain%s", path);
  }
  const long fyl = fx0 ? sh.yl : s.yl;
  int rc = GT_OK;
  auto dimid = [&](const char* nm) -> uint32_t {
    for (size_t q = 0; q < dims.size(); q++)
      if (dims[q].first == nm) return (uint32_t)q;
    return 0xFFFFFFFFu;
  };
  const uint32_t dx = dimid("xax"), dy = dimid("yax"), dz = dimid("zax");
  // the scalars open_restart_ncdf.F90:48-62 requires and read_restart_ncdf
  // reads first (:71-81)
  gt_restart_time got = {0, 0, 0, 0.0};
  const char* sname[4] = {"loop", "julianday", "secondsofday", "timestep"};
  for (int q = 0; q < 4 && rc == GT_OK; q++) {
    auto it = vars.find(sname[q]);
    const uint32_t type = q < 3 ? NC_INT : NC_DOUBLE;
    uint8_t b[8];
    const size_t sz = q < 3 ? 4 : 8;
    if (it == vars.end() || it->second.type != type || !it->second.dims.empty() ||
        std::fseek(fp, (long)it->second.begin, SEEK_SET) != 0 || std::fread(b, 1, sz, fp) != sz) {
      rc = fail(GT_EARG, "gt_restart_read: %s: no scalar '%s' of the reference's type", path,
                sname[q]);
      break;
    }
    if (q < 3) {
      const int32_t v = (int32_t)((uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 |
                                  (uint32_t)b[2] << 8 | b[3]);
      (q == 0 ? got.loop : q == 1 ? got.julianday : got.secondsofday) = v;
    } else {
      get_be_doubles(&got.timestep, b, 1);
    }
  }
  if (rc == GT_OK && tm) *tm = got;
  std::vector<uint8_t> row((size_t)s.xl * 8);
  for (int v = 0; v < n && rc == GT_OK; v++) {
    auto it = vars.find(f[v].name ? f[v].name : "");
    const std::vector<uint32_t> want3 = {dz, dy, dx}, want2 = {dy, dx};
    if (it == vars.end() || it->second.type != NC_DOUBLE ||
        it->second.dims != (f[v].ndim == 3 ? want3 : want2)) {
      rc = fail(GT_EARG, "gt_restart_read: %s: no FP64 variable %s with the field's dimensions",
                path, f[v].name ? f[v].name : "?");
      break;
    }
    const long nk = f[v].ndim == 3 ? s.NK : 1;
    for (long k = 0; k < nk && rc == GT_OK; k++)
      for (long j = 0; j < s.yl; j++) {
        const long at = (long)it->second.begin + 8 * ((k * fyl + fx0 + j) * fxl + fx0);
        if (std::fseek(fp, at, SEEK_SET) != 0 ||
            std::fread(row.data(), 1, row.size(), fp) != row.size()) {
          rc = fail(GT_EARG, "gt_restart_read: %s: short read of %s", path, f[v].name);
          break;
        }
        double* dst = f[v].data + (k * s.NY + (s.y0 + j)) * s.NX + s.x0;
        get_be_doubles(dst, row.data(), (size_t)s.xl);
      }
  }
  std::fclose(fp);
  return rc;
}
