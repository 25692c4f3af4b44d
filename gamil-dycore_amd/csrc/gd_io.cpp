// gd_io.cpp -- history and restart files (history_mod.F90, restart_mod.F90,
// io_mod.F90) written and read in the NetCDF classic format (CDF-1: big-
// endian, 32-bit offsets, one unlimited "time" dimension), which is what the
// reference creates with NF90_CREATE(path, NF90_CLOBBER).  The format is
// implemented here from its public specification; no NetCDF library is used.
//
// Host-only code: the caller passes host arrays (e.g. from gd_get_state and
// gd_get_diag_fields), one output time per file.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gamil_dycore.h"

namespace {

constexpr uint32_t kNcDim = 0x0A, kNcVar = 0x0B, kNcAtt = 0x0C;
constexpr uint32_t kChar = 2, kInt = 4, kDouble = 6;
constexpr double kFillDouble = 9.9692099683868690e+36;  // NC_FILL_DOUBLE
constexpr double kPi = 3.14159265358979323846;  // = atan(1d0)*4 in FP64

thread_local std::string g_io_error;

int io_fail(const std::string &m) {
  g_io_error = m;
  return GD_ERR_INVALID_ARG;
}

void put32(std::vector<unsigned char> &b, uint32_t v) {
  for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
}
void put_double(std::vector<unsigned char> &b, double d) {
  uint64_t v;
  std::memcpy(&v, &d, 8);
  for (int k = 7; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
}
void pad4(std::vector<unsigned char> &b) {
  while (b.size() % 4) b.push_back(0);
}
void put_name(std::vector<unsigned char> &b, const std::string &s) {
  put32(b, (uint32_t)s.size());
  b.insert(b.end(), s.begin(), s.end());
  pad4(b);
}

struct Att {
  std::string name;
  uint32_t type;
  std::string text;
  std::vector<double> dv;
  std::vector<int32_t> iv;
};

Att att_text(const std::string &n, const std::string &t) { return Att{n, kChar, t, {}, {}}; }
Att att_double(const std::string &n, double v) { return Att{n, kDouble, "", {v}, {}}; }
Att att_int(const std::string &n, int32_t v) { return Att{n, kInt, "", {}, {v}}; }

void put_atts(std::vector<unsigned char> &b, const std::vector<Att> &atts) {
  if (atts.empty()) {
    put32(b, 0);
    put32(b, 0);
    return;
  }
  put32(b, kNcAtt);
  put32(b, (uint32_t)atts.size());
  for (const Att &a : atts) {
    put_name(b, a.name);
    put32(b, a.type);
    if (a.type == kChar) {
      put32(b, (uint32_t)a.text.size());
      b.insert(b.end(), a.text.begin(), a.text.end());
      pad4(b);
    } else if (a.type == kDouble) {
      put32(b, (uint32_t)a.dv.size());
      for (double d : a.dv) put_double(b, d);
    } else {
      put32(b, (uint32_t)a.iv.size());
      for (int32_t v : a.iv) put32(b, (uint32_t)v);
    }
  }
}

struct Var {
  std::string name;
  std::vector<int> dims;          // indices into the dimension list, slowest first
  std::vector<Att> atts;
  std::vector<double> data;       // values of one record (record variables) or all
};

// A one-record CDF-1 file: dims[0] is the unlimited dimension.
struct NcFile {
  std::vector<std::pair<std::string, uint32_t>> dims;  // length 0 = unlimited
  std::vector<Att> gatts;
  std::vector<Var> vars;

  bool is_record(const Var &v) const { return !v.dims.empty() && dims[v.dims[0]].second == 0; }
  size_t count(const Var &v) const {
    size_t n = 1;
    for (int d : v.dims)
      if (dims[d].second) n *= dims[d].second;
    return n;
  }

  int write(const char *path) const {
    // header with placeholder offsets, sized first
    auto header = [&](const std::vector<uint32_t> &begins) {
      std::vector<unsigned char> b;
      b.push_back('C'); b.push_back('D'); b.push_back('F'); b.push_back(1);
      put32(b, 1);  // numrecs
      put32(b, kNcDim);
      put32(b, (uint32_t)dims.size());
      for (const auto &d : dims) {
        put_name(b, d.first);
        put32(b, d.second);
      }
      put_atts(b, gatts);
      put32(b, kNcVar);
      put32(b, (uint32_t)vars.size());
      for (size_t k = 0; k < vars.size(); ++k) {
        const Var &v = vars[k];
        put_name(b, v.name);
        put32(b, (uint32_t)v.dims.size());
        for (int d : v.dims) put32(b, (uint32_t)d);
        put_atts(b, v.atts);
        put32(b, kDouble);
        put32(b, (uint32_t)(count(v) * 8));
        put32(b, begins[k]);
      }
      return b;
    };
    std::vector<uint32_t> begins(vars.size(), 0);
    const size_t hsize = header(begins).size();
    // non-record variables first, then the single record
    size_t off = hsize;
    for (size_t k = 0; k < vars.size(); ++k)
      if (!is_record(vars[k])) {
        begins[k] = (uint32_t)off;
        off += count(vars[k]) * 8;
      }
    for (size_t k = 0; k < vars.size(); ++k)
      if (is_record(vars[k])) {
        begins[k] = (uint32_t)off;
        off += count(vars[k]) * 8;
      }
    if (off > 0x7fffffffu) return io_fail("file too large for the classic NetCDF format");
    std::vector<unsigned char> b = header(begins);
    for (int pass = 0; pass < 2; ++pass)
      for (const Var &v : vars)
        if (is_record(v) == (pass == 1))
          for (double d : v.data) put_double(b, d);
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(std::string("cannot create ") + path);
    const size_t w = std::fwrite(b.data(), 1, b.size(), f);
    const int c = std::fclose(f);
    if (w != b.size() || c != 0) return io_fail(std::string("cannot write ") + path);
    return GD_OK;
  }
};

// lon, lat, ilon, ilat axes in degrees, evaluated as mesh_mod.F90:70-85:
// full points first, half points = full + dlon/2 (dlat/2), then times
// rad_to_deg = 180/pi (params_mod.F90:7); the north pole row is set to 90
void add_axes(NcFile &f, int N, int M) {
  const double rad_to_deg = 180.0 / kPi;
  const double dlon = 2 * kPi / N, dlat = kPi / (M - 1);
  std::vector<double> lon(N), ilon(N), lat(M), ilat(M - 1);
  for (int i = 0; i < N; ++i) {
    const double fl = i * dlon;
    lon[i] = fl * rad_to_deg;
    ilon[i] = (fl + 0.5 * dlon) * rad_to_deg;
  }
  for (int j = 0; j < M - 1; ++j) {
    const double fl = -0.5 * kPi + j * dlat;
    lat[j] = fl * rad_to_deg;
    ilat[j] = (fl + 0.5 * dlat) * rad_to_deg;
  }
  lat[M - 1] = 90.0;
  f.vars.push_back(Var{"lon", {1}, {att_text("long_name", "Longitude"), att_text("units", "degrees_east")}, lon});
  f.vars.push_back(Var{"lat", {2}, {att_text("long_name", "Latitude"), att_text("units", "degrees_north")}, lat});
  f.vars.push_back(Var{"ilon", {3}, {att_text("long_name", "Longitude"), att_text("units", "degrees_east")}, ilon});
  f.vars.push_back(Var{"ilat", {4}, {att_text("long_name", "Latitude"), att_text("units", "degrees_north")}, ilat});
}

void base_file(NcFile &f, const gd_io_info *in, const char *dataset, int N, int M) {
  f.dims = {{"time", 0}, {"lon", (uint32_t)N}, {"lat", (uint32_t)M}, {"ilon", (uint32_t)N}, {"ilat", (uint32_t)(M - 1)}};
  f.gatts.push_back(att_text("dataset", dataset));
  f.gatts.push_back(att_text("desc", in->desc ? in->desc : ""));
  f.gatts.push_back(att_text("author", in->author ? in->author : ""));
  const std::string units = std::string(in->time_units ? in->time_units : "days") + " since " +
                            (in->start_time ? in->start_time : "");
  f.vars.push_back(Var{"time", {0}, {att_text("long_name", "Time"), att_text("units", units)}, {in->time_value}});
  add_axes(f, N, M);
}

// rows x N host block from a strided array (row r starts at p + r*ld)
std::vector<double> block(const double *p, int rows, int N, int64_t ld) {
  std::vector<double> out((size_t)rows * N);
  for (int r = 0; r < rows; ++r)
    for (int i = 0; i < N; ++i) out[(size_t)r * N + i] = p[(size_t)r * ld + i];
  return out;
}

Var field(const std::string &name, const char *long_name, const char *units, std::vector<int> dims,
          std::vector<double> data) {
  return Var{name, std::move(dims), {att_text("long_name", long_name), att_text("units", units)}, std::move(data)};
}

// ---- reader (classic CDF-1 and 64-bit-offset CDF-2) ----

size_t type_size(uint32_t t) { return t == 1 || t == 2 ? 1 : t == 3 ? 2 : t == 6 ? 8 : 4; }

struct Reader {
  std::vector<unsigned char> b;
  size_t p = 0;
  bool ok = true;
  int version = 1;
  uint32_t u32() {
    if (p + 4 > b.size()) { ok = false; return 0; }
    uint32_t v = ((uint32_t)b[p] << 24) | ((uint32_t)b[p + 1] << 16) | ((uint32_t)b[p + 2] << 8) | b[p + 3];
    p += 4;
    return v;
  }
  uint64_t off() { return version == 2 ? ((uint64_t)u32() << 32) | u32() : u32(); }
  std::string name() {
    const uint32_t n = u32();
    if (p + n > b.size()) { ok = false; return ""; }
    std::string s((const char *)&b[p], n);
    p += (n + 3) & ~3u;
    return s;
  }
  double dbl(size_t at) const {
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) v = (v << 8) | b[at + k];
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
  // one numeric value of NetCDF type t (byte, short, int, float, double)
  double num(uint32_t t, size_t at) const {
    if (t == 6) return dbl(at);
    if (t == 1) return (double)(int8_t)b[at];
    if (t == 3) return (double)(int16_t)(((uint16_t)b[at] << 8) | b[at + 1]);
    const uint32_t v = ((uint32_t)b[at] << 24) | ((uint32_t)b[at + 1] << 16) | ((uint32_t)b[at + 2] << 8) | b[at + 3];
    if (t == 4) return (double)(int32_t)v;
    float f;
    std::memcpy(&f, &v, 4);
    return (double)f;
  }
};

struct RVar {
  std::string name;
  std::vector<uint32_t> dimids;
  uint32_t type = 0;
  uint64_t begin = 0;
};

struct RAtt {
  std::string name;
  uint32_t type;
  std::string text;
  std::vector<double> dv;
};

bool read_atts(Reader &r, std::vector<RAtt> *out) {
  const uint32_t tag = r.u32(), n = r.u32();
  if (tag == 0) return r.ok;
  if (tag != kNcAtt) return false;
  for (uint32_t k = 0; k < n && r.ok; ++k) {
    RAtt a;
    a.name = r.name();
    a.type = r.u32();
    const uint32_t cnt = r.u32();
    const size_t bytes = cnt * type_size(a.type);
    if (r.p + bytes > r.b.size()) return false;
    if (a.type == kChar) a.text.assign((const char *)&r.b[r.p], cnt);
    else
      for (uint32_t q = 0; q < cnt; ++q) a.dv.push_back(r.num(a.type, r.p + q * type_size(a.type)));
    r.p += (bytes + 3) & ~(size_t)3;
    if (out) out->push_back(a);
  }
  return r.ok;
}

}  // namespace

extern "C" {

const char *gd_io_last_error(void) { return g_io_error.c_str(); }

int gd_history_write(const char *path, const gd_io_info *in, int32_t N, int32_t M, const double *u,
                     const double *v, const double *gd, const double *ghs, const double *div,
                     const double *vor, int64_t ld, double total_energy, double total_mass) {
  if (!path || !in || N < 1 || M < 3 || ld < N || !u || !v || !gd || !ghs || !div || !vor)
    return io_fail("bad history arguments");
  NcFile f;
  base_file(f, in, "hist0", N, M);
  f.gatts.push_back(att_double("time_step_size", in->time_step_size));
  f.gatts.push_back(att_text("time_scheme", in->time_scheme ? in->time_scheme : ""));
  f.gatts.push_back(att_text("split_scheme", in->split_scheme ? in->split_scheme : ""));
  f.gatts.push_back(att_int("subcycles", in->subcycles));
  // C grid -> A grid (history_mod.F90:100-106): u from the two neighbouring
  // half-longitude points (periodic), v from the two neighbouring half
  // latitudes (zero beyond the last half row, as the reference's halo)
  std::vector<double> ua((size_t)M * N), va((size_t)M * N, kFillDouble), gh((size_t)M * N);
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < N; ++i) {
      const int im = i == 0 ? N - 1 : i - 1;
      ua[(size_t)j * N + i] = 0.5 * (u[(size_t)j * ld + i] + u[(size_t)j * ld + im]);
      gh[(size_t)j * N + i] = gd[(size_t)j * ld + i] + ghs[(size_t)j * ld + i];
    }
  // the reference writes rows 1..M-1 of its A-grid v into the M-row "lat"
  // dimension (history_mod.F90:115); the last row keeps the fill value
  for (int j = 0; j < M - 1; ++j)
    for (int i = 0; i < N; ++i) {
      const double vj = v[(size_t)j * ld + i];
      const double vm = j >= 1 ? v[(size_t)(j - 1) * ld + i] : 0.0;
      va[(size_t)j * N + i] = 0.5 * (vj + vm);
    }
  f.vars.push_back(field("u", "u wind component", "m s-1", {0, 2, 1}, ua));
  f.vars.push_back(field("v", "v wind component", "m s-1", {0, 2, 1}, va));
  f.vars.push_back(field("gh", "geopotential height", "m2 s-2", {0, 2, 1}, gh));
  f.vars.push_back(field("ghs", "surface geopotential", "m2 s-2", {0, 2, 1}, block(ghs, M, N, ld)));
  // vor lives on (ilon, ilat) and div on (lon, lat); the reference declares
  // them the other way round (history_mod.F90:48-49 vs 117-118)
  f.vars.push_back(field("vor", "relative vorticity", "s-1", {0, 4, 3}, block(vor, M - 1, N, ld)));
  f.vars.push_back(field("div", "divergence", "s-1", {0, 2, 1}, block(div, M, N, ld)));
  f.vars.push_back(field("te", "total energy", "m4 s-4", {0}, {total_energy}));
  f.vars.push_back(field("tm", "total mass", "m2 s-2", {0}, {total_mass}));
  return f.write(path);
}

int gd_restart_write(const char *path, const gd_io_info *in, int32_t N, int32_t M, const double *u,
                     const double *v, const double *gd, const double *ghs, int64_t ld) {
  if (!path || !in || N < 1 || M < 3 || ld < N || !u || !v || !gd || !ghs)
    return io_fail("bad restart arguments");
  NcFile f;
  base_file(f, in, "restart", N, M);
  f.gatts.push_back(att_text("restart_time", in->restart_time ? in->restart_time : ""));
  f.gatts.push_back(att_double("elapsed_seconds", in->elapsed_seconds));
  f.vars.push_back(field("u", "u wind component", "m s-1", {0, 2, 3}, block(u, M, N, ld)));
  f.vars.push_back(field("v", "v wind component", "m s-1", {0, 4, 1}, block(v, M - 1, N, ld)));
  f.vars.push_back(field("gd", "geopotential depth", "m2 s-2", {0, 2, 1}, block(gd, M, N, ld)));
  f.vars.push_back(field("ghs", "surface geopotential", "m2 s-2", {0, 2, 1}, block(ghs, M, N, ld)));
  return f.write(path);
}

int gd_restart_read(const char *path, int32_t N, int32_t M, double *u, double *v, double *gd, double *ghs,
                    int64_t ld, char *restart_time, int32_t cap, double *elapsed_seconds) {
  if (!path || N < 1 || M < 3 || ld < N) return io_fail("bad restart arguments");
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return io_fail(std::string("cannot open ") + path);
  Reader r;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) r.b.insert(r.b.end(), buf, buf + n);
  std::fclose(fp);
  if (r.b.size() < 8 || r.b[0] != 'C' || r.b[1] != 'D' || r.b[2] != 'F' || (r.b[3] != 1 && r.b[3] != 2))
    return io_fail(std::string(path) + " is not a classic NetCDF file");
  r.version = r.b[3];
  r.p = 4;
  const uint32_t numrecs = r.u32();
  std::vector<std::pair<std::string, uint32_t>> dims;
  if (r.u32() == kNcDim) {
    const uint32_t nd = r.u32();
    for (uint32_t k = 0; k < nd && r.ok; ++k) {
      std::string nm = r.name();
      dims.push_back({nm, r.u32()});
    }
  } else {
    r.u32();
  }
  std::vector<RAtt> gatts;
  if (!read_atts(r, &gatts)) return io_fail(std::string(path) + ": bad global attributes");
  std::vector<RVar> vars;
  const uint32_t vt = r.u32(), nv = r.u32();
  if (vt == kNcVar)
    for (uint32_t k = 0; k < nv && r.ok; ++k) {
      RVar var;
      var.name = r.name();
      const uint32_t nd = r.u32();
      for (uint32_t q = 0; q < nd && r.ok; ++q) var.dimids.push_back(r.u32());
      if (!read_atts(r, nullptr)) return io_fail(std::string(path) + ": bad variable attributes");
      var.type = r.u32();
      r.u32();  // vsize
      var.begin = r.off();
      vars.push_back(var);
    }
  if (!r.ok) return io_fail(std::string(path) + ": truncated header");
  if (numrecs < 1) return io_fail(std::string(path) + ": no record");
  // record size (one record of every record variable)
  uint64_t recsize = 0;
  auto is_rec = [&](const RVar &v) { return !v.dimids.empty() && v.dimids[0] < dims.size() && dims[v.dimids[0]].second == 0; };
  auto count = [&](const RVar &v) {
    uint64_t c = 1;
    for (uint32_t d : v.dimids)
      if (d < dims.size() && dims[d].second) c *= dims[d].second;
    return c;
  };
  for (const RVar &v : vars)
    if (is_rec(v)) recsize += (count(v) * type_size(v.type) + 3) & ~(uint64_t)3;
  auto get = [&](const char *name, int rows, double *dst) -> int {
    if (!dst) return GD_OK;
    for (const RVar &v : vars) {
      if (v.name != name) continue;
      if (v.type != kDouble) return io_fail(std::string(name) + " is not a double variable");
      if (count(v) != (uint64_t)rows * N) return io_fail(std::string(name) + " has the wrong size");
      // the last record
      const uint64_t at = v.begin + (is_rec(v) ? (uint64_t)(numrecs - 1) * recsize : 0);
      if (at > r.b.size() || count(v) * 8 > r.b.size() - at) return io_fail(std::string(path) + ": truncated data");
      for (int j = 0; j < rows; ++j)
        for (int i = 0; i < N; ++i) dst[(size_t)j * ld + i] = r.dbl(at + ((size_t)j * N + i) * 8);
      return GD_OK;
    }
    return io_fail(std::string(path) + ": no variable " + name);
  };
  int rc;
  if ((rc = get("u", M, u)) || (rc = get("v", M - 1, v)) || (rc = get("gd", M, gd)) || (rc = get("ghs", M, ghs)))
    return rc;
  for (const RAtt &a : gatts) {
    if (a.name == "restart_time" && restart_time && cap > 0) std::snprintf(restart_time, cap, "%s", a.text.c_str());
    if (a.name == "elapsed_seconds" && elapsed_seconds && !a.dv.empty()) *elapsed_seconds = a.dv[0];
  }
  return GD_OK;
}

}  // extern "C"
