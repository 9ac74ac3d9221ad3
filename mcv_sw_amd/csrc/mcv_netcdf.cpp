// History file of the reference's output_mod.F90 in the netCDF classic
// "64-bit offset" format (CDF-2), written from the published format
// specification (NetCDF Classic Format Specification, Unidata) without
// libnetcdf.  Host code only.
//
//   mcv_history_init   history_init (output_mod.F90:13-166): creates the file
//                      with the reference's dimensions, variables and
//                      attributes, and writes the static variables.
//   mcv_history_write  history_write_stat (output_mod.F90:168-223): writes one
//                      record (time slot) of the 12 record variables.
//
// Layout (all values big-endian):
//   header   "CDF\x02", numrecs, dimension list, global attributes, variables
//            (name, dimension ids, attributes, type, vsize, 8-byte begin)
//   data     the 6 static variables one after the other, then the records:
//            record k holds every record variable's slab for time slot k, in
//            definition order (time, uP, vP, ..., phiC).
// Fortran arrays (i, j, panel) are C arrays [panel][j][i], so a variable's
// Fortran dimensions appear reversed in the header and its data is the
// Fortran array in memory order.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mcv_hip.h"

namespace mcv {
int fail(int code, const char *fmt, ...);
int history_write_checked(const char *path, int32_t time_slot, const mcv_history_fields *rec,
                          long long npoint, long long ncell);
}

namespace {

using mcv::fail;

enum : uint32_t { kNcChar = 2, kNcInt = 4, kNcDouble = 6 };
enum : uint32_t { kTagDim = 10, kTagVar = 11, kTagAtt = 12 };

// constants_mod.F90:13,17,20
const double kPi = 2.0 * std::asin(1.0);
const double kR2D = 180.0 / kPi;
const double kFillValue = -9999999999999999.0;

struct Att {
  std::string name;
  uint32_t type;
  std::string chars;           // kNcChar
  std::vector<int32_t> ints;   // kNcInt
  std::vector<double> dbls;    // kNcDouble
};

Att att_text(const char *name, const char *v) { return Att{name, kNcChar, v, {}, {}}; }
Att att_int(const char *name, int32_t v) { return Att{name, kNcInt, "", {v}, {}}; }
Att att_dbl(const char *name, double v) { return Att{name, kNcDouble, "", {}, {v}}; }

struct Var {
  std::string name;
  uint32_t type;
  std::vector<uint32_t> dims;  // C order (slowest first)
  std::vector<Att> atts;
  uint64_t vsize = 0;          // bytes per variable (per record for record variables)
  uint64_t begin = 0;
  bool record = false;
};

struct Header {
  std::vector<std::pair<std::string, uint32_t>> dims;  // (name, length); 0 = record dim
  std::vector<Att> gatts;
  std::vector<Var> vars;
  uint32_t numrecs = 0;
  uint64_t data_begin = 0, rec_begin = 0, recsize = 0;
};

// ---- encoding -------------------------------------------------------------
struct Out {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void att(const Att &a) {
    name(a.name);
    u32(a.type);
    if (a.type == kNcChar) {
      u32((uint32_t)a.chars.size());
      b.insert(b.end(), a.chars.begin(), a.chars.end());
    } else if (a.type == kNcInt) {
      u32((uint32_t)a.ints.size());
      for (int32_t v : a.ints) u32((uint32_t)v);
    } else {
      u32((uint32_t)a.dbls.size());
      for (double v : a.dbls) {
        uint64_t bits;
        std::memcpy(&bits, &v, 8);
        u64(bits);
      }
    }
    pad();
  }
  void atts(const std::vector<Att> &l) {
    if (l.empty()) { u32(0); u32(0); return; }
    u32(kTagAtt);
    u32((uint32_t)l.size());
    for (const Att &a : l) att(a);
  }
};

std::vector<uint8_t> encode_header(const Header &h) {
  Out o;
  o.b = {'C', 'D', 'F', 2};
  o.u32(h.numrecs);
  o.u32(kTagDim);
  o.u32((uint32_t)h.dims.size());
  for (const auto &d : h.dims) { o.name(d.first); o.u32(d.second); }
  o.atts(h.gatts);
  o.u32(kTagVar);
  o.u32((uint32_t)h.vars.size());
  for (const Var &v : h.vars) {
    o.name(v.name);
    o.u32((uint32_t)v.dims.size());
    for (uint32_t d : v.dims) o.u32(d);
    o.atts(v.atts);
    o.u32(v.type);
    o.u32(v.vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)v.vsize);
    o.u64(v.begin);
  }
  return o.b;
}

uint32_t type_size(uint32_t t) { return t == kNcDouble ? 8 : t == kNcInt ? 4 : 1; }

// Sizes and offsets: static variables back to back after the header, then
// the record section.
void lay_out(Header &h) {
  for (Var &v : h.vars) {
    uint64_t n = type_size(v.type);
    v.record = !v.dims.empty() && h.dims[v.dims[0]].second == 0;
    for (size_t k = v.record ? 1 : 0; k < v.dims.size(); ++k) n *= h.dims[v.dims[k]].second;
    v.vsize = (n + 3) / 4 * 4;
  }
  uint64_t off = encode_header(h).size();  // begins have fixed width: size is final
  h.data_begin = off;
  for (Var &v : h.vars)
    if (!v.record) { v.begin = off; off += v.vsize; }
  h.rec_begin = off;
  h.recsize = 0;
  for (Var &v : h.vars)
    if (v.record) { v.begin = off + h.recsize; h.recsize += v.vsize; }
}

// ---- decoding (files this library wrote, or any CDF-1/CDF-2 file) ---------
struct In {
  const std::vector<uint8_t> &b;
  size_t p = 0;
  bool ok = true;
  uint32_t u32() {
    if (p + 4 > b.size()) { ok = false; return 0; }
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) v = v << 8 | b[p++];
    return v;
  }
  uint64_t u64() { uint64_t hi = u32(); return hi << 32 | u32(); }
  void skip(size_t n) { p += (n + 3) / 4 * 4; if (p > b.size()) ok = false; }
  std::string name() {
    uint32_t n = u32();
    if (!ok || p + n > b.size()) { ok = false; return ""; }
    std::string s(b.begin() + p, b.begin() + p + n);
    skip(n);
    return s;
  }
  void skip_atts() {
    uint32_t tag = u32(), n = u32();
    if (tag == 0) return;
    for (uint32_t k = 0; k < n && ok; ++k) {
      name();
      uint32_t t = u32(), cnt = u32();
      skip((size_t)cnt * type_size(t));
    }
  }
};

int parse_header(const std::vector<uint8_t> &b, int version, Header &h) {
  In in{b};
  in.p = 4;
  h.numrecs = in.u32();
  uint32_t tag = in.u32(), n = in.u32();
  if (tag != kTagDim && !(tag == 0 && n == 0)) return -1;
  for (uint32_t k = 0; k < n && in.ok; ++k) {
    std::string nm = in.name();
    h.dims.emplace_back(nm, in.u32());
  }
  in.skip_atts();
  tag = in.u32();
  n = in.u32();
  if (tag != kTagVar && !(tag == 0 && n == 0)) return -1;
  h.recsize = 0;
  for (uint32_t k = 0; k < n && in.ok; ++k) {
    Var v;
    v.name = in.name();
    uint32_t nd = in.u32();
    for (uint32_t d = 0; d < nd && in.ok; ++d) v.dims.push_back(in.u32());
    in.skip_atts();
    v.type = in.u32();
    v.vsize = in.u32();
    v.begin = version == 2 ? in.u64() : in.u32();
    for (uint32_t d : v.dims)
      if (d >= h.dims.size()) return -1;
    v.record = !v.dims.empty() && h.dims[v.dims[0]].second == 0;
    if (v.record) h.recsize += v.vsize;
    h.vars.push_back(v);
  }
  return in.ok ? 0 : -1;
}

// ---- file I/O -------------------------------------------------------------
void put_be64(uint8_t *dst, const double *src, size_t n, double scale, bool scaled) {
  for (size_t k = 0; k < n; ++k) {
    const double v = scaled ? src[k] * scale : src[k];
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    for (int s = 0; s < 8; ++s) dst[8 * k + s] = (uint8_t)(bits >> (56 - 8 * s));
  }
}

int write_at(FILE *f, uint64_t off, const void *p, size_t n, const char *path) {
  if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fwrite(p, 1, n, f) != n)
    return fail(MCV_EINVAL, "write to %s failed", path);
  return MCV_OK;
}

// Writes n doubles (times `scale` when `scaled`) big-endian at `off`, in chunks.
int write_doubles(FILE *f, uint64_t off, const double *src, size_t n, const char *path,
                  double scale = 1.0, bool scaled = false) {
  std::vector<uint8_t> buf;
  const size_t chunk = 1 << 20;
  for (size_t k = 0; k < n; k += chunk) {
    const size_t m = std::min(chunk, n - k);
    buf.resize(8 * m);
    put_be64(buf.data(), src + k, m, scale, scaled);
    int rc = write_at(f, off + 8 * k, buf.data(), buf.size(), path);
    if (rc) return rc;
  }
  return MCV_OK;
}

const char *const kRecordNames[12] = {"time", "uP", "vP", "contraUP", "contraVP", "phiP",
                                      "phiP_t", "zonal_wind", "meridional_wind", "uC", "vC",
                                      "phiC"};

}  // namespace

extern "C" {

int mcv_history_init(const char *path, const mcv_history_info *info, const double *lonC,
                     const double *latC, const double *lonP, const double *latP,
                     const double *areaCell, const double *phi_s) {
  if (!path || !info || !lonC || !latC || !lonP || !latP || !areaCell || !phi_s)
    return fail(MCV_EINVAL, "mcv_history_init: NULL argument");
  if (info->dof != 3 && info->dof != 4) return fail(MCV_EINVAL, "dof must be 3 or 4");
  if (info->n < 1) return fail(MCV_EINVAL, "n must be >= 1");
  if (info->xhalo != 1 || info->yhalo != 1)
    return fail(MCV_EINVAL, "the library's arrays carry one halo cell (xhalo = yhalo = 1)");
  const int H = info->dof - 1, n = info->n, npv = n * H + 1;
  const int xh = info->xhalo, yh = info->yhalo;
  // parameters_mod.F90:150-199
  const int ics = 1 - xh, ice = n + xh, jcs = 1 - yh, jce = n + yh;
  const int ips = 1 - xh * H, ipe = npv + xh * H, jps = 1 - yh * H, jpe = npv + yh * H;
  const uint32_t nxc = ice - ics + 1, nyc = jce - jcs + 1, nxp = ipe - ips + 1,
                 nyp = jpe - jps + 1;

  Header h;
  h.dims = {{"lonC", nxc}, {"latC", nyc}, {"lonP", nxp}, {"latP", nyp}, {"nPatch", 6},
            {"time", 0}};
  enum { dLonC, dLatC, dLonP, dLatP, dPatch, dTime };
  h.gatts = {att_int("MCV_ORDER", info->dof), att_dbl("dx", info->dx * kR2D),
             att_dbl("dy", info->dy * kR2D), att_dbl("dt", info->dt), att_int("xhalo", xh),
             att_int("yhalo", yh), att_int("case_num", info->case_num), att_int("ics", ics),
             att_int("ice", ice), att_int("jcs", jcs), att_int("jce", jce), att_int("ips", ips),
             att_int("ipe", ipe), att_int("jps", jps), att_int("jpe", jpe), att_int("ids", 1),
             att_int("ide", npv), att_int("jds", 1), att_int("jde", npv), att_int("ifs", 1),
             att_int("ife", 6)};
  const std::vector<uint32_t> cell = {dPatch, dLatC, dLonC}, point = {dPatch, dLatP, dLonP};
  const std::vector<uint32_t> cellt = {dTime, dPatch, dLatC, dLonC},
                              pointt = {dTime, dPatch, dLatP, dLonP};
  auto var = [&](const char *name, uint32_t type, const std::vector<uint32_t> &dims,
                 const char *units, const char *long_name, bool fill) {
    Var v;
    v.name = name;
    v.type = type;
    v.dims = dims;
    v.atts = {att_text("units", units), att_text("long_name", long_name)};
    if (fill) v.atts.push_back(att_dbl("_FillValue", kFillValue));
    h.vars.push_back(v);
  };
  // output_mod.F90:47-137 (definition order, units, long names and fill values)
  var("lonC", kNcDouble, cell, "degree_east", "longitude on sphere coordinate for Cells", false);
  var("latC", kNcDouble, cell, "degree_north", "latitude on sphere coordinate for Cells", false);
  var("lonP", kNcDouble, point, "degree_east", "longitude on sphere coordinate for Points", false);
  var("latP", kNcDouble, point, "degree_north", "latitude on sphere coordinate for Points", false);
  var("areaCell", kNcDouble, cell, "m^2", "area of cells", false);
  var("phiP_s", kNcDouble, point, "m^2/s^2", "surface height on points", true);
  var("time", kNcInt, {dTime}, "seconds", "time", false);
  var("uP", kNcDouble, pointt, "m/s", "covariant wind vector on x direcvtion", true);
  var("vP", kNcDouble, pointt, "m/s", "covariant wind vector on y direcvtion", true);
  var("contraUP", kNcDouble, pointt, "m/s", "contravariant wind vector on x direcvtion", true);
  var("contraVP", kNcDouble, pointt, "m/s", "contravariant wind vector on y direcvtion", true);
  var("phiP", kNcDouble, pointt, "m^2/s^2", "geopotential height on points", true);
  var("phiP_t", kNcDouble, pointt, "m^2/s^2", "total geopotential height on points", true);
  var("zonal_wind", kNcDouble, pointt, "m/s", "zonal wind", true);
  var("meridional_wind", kNcDouble, pointt, "m/s", "meridional wind", true);
  var("uC", kNcDouble, cellt, "m/s", "covariant wind VIA on x direcvtion", true);
  var("vC", kNcDouble, cellt, "m/s", "covariant wind VIA on y direcvtion", true);
  var("phiC", kNcDouble, cellt, "m^2/s^2", "geopotential height VIA", true);
  lay_out(h);

  FILE *f = std::fopen(path, "wb");  // NF90_CLOBBER
  if (!f) return fail(MCV_EINVAL, "cannot create %s", path);
  const std::vector<uint8_t> hdr = encode_header(h);
  int rc = write_at(f, 0, hdr.data(), hdr.size(), path);
  // output_mod.F90:159-164: coordinates in degrees, area, surface geopotential
  const double *src[6] = {lonC, latC, lonP, latP, areaCell, phi_s};
  const bool deg[6] = {true, true, true, true, false, false};
  for (int k = 0; k < 6 && !rc; ++k)
    rc = write_doubles(f, h.vars[k].begin, src[k], h.vars[k].vsize / 8, path, kR2D, deg[k]);
  if (std::fclose(f) != 0 && !rc) rc = fail(MCV_EINVAL, "closing %s failed", path);
  return rc;
}

int mcv_history_write(const char *path, int32_t time_slot, int32_t dof, int32_t n,
                      const mcv_history_fields *rec) {
  // the caller's arrays: (nh, nh, 6) points, (n+2, n+2, 6) cells
  if ((dof != 3 && dof != 4) || n < 1 || n > 4096)
    return mcv::fail(MCV_EINVAL, "mcv_history_write: dof must be 3 or 4 and n in [1, 4096], got "
                                 "%d, %d", dof, n);
  const long long nh = (long long)n * (dof - 1) + 1 + 2 * (dof - 1);
  return mcv::history_write_checked(path, time_slot, rec, 6 * nh * nh,
                                    6 * (long long)(n + 2) * (n + 2));
}

}  // extern "C"

namespace mcv {

// mcv_history_write; with npoint, ncell > 0 the file's record variables must
// hold exactly that many values per record (the caller's array sizes).
int history_write_checked(const char *path, int32_t time_slot, const mcv_history_fields *rec,
                          long long npoint, long long ncell) {
  if (!path || !rec) return fail(MCV_EINVAL, "mcv_history_write: NULL argument");
  if (time_slot < 1) return fail(MCV_EINVAL, "time_slot must be >= 1, got %d", time_slot);
  const double *src[11] = {rec->u, rec->v, rec->contraU, rec->contraV, rec->phi, rec->phi_t,
                           rec->zonal, rec->meridional, rec->uC, rec->vC, rec->phiC};
  for (const double *p : src)
    if (!p) return fail(MCV_EINVAL, "mcv_history_write: NULL field");
  FILE *f = std::fopen(path, "r+b");  // NF90_WRITE
  if (!f) return fail(MCV_EINVAL, "cannot open %s", path);
  // the header: read until it parses (it is small)
  std::vector<uint8_t> b(4096);
  Header h;
  int version = 0, rc = MCV_OK;
  for (;;) {
    std::fseek(f, 0, SEEK_SET);
    const size_t got = std::fread(b.data(), 1, b.size(), f);
    if (got < 8 || std::memcmp(b.data(), "CDF", 3) != 0 || (b[3] != 1 && b[3] != 2)) {
      rc = fail(MCV_EINVAL, "%s is not a netCDF classic file", path);
      break;
    }
    version = b[3];
    std::vector<uint8_t> part(b.begin(), b.begin() + got);
    h = Header{};
    if (parse_header(part, version, h) == 0) break;
    if (got < b.size()) {
      rc = fail(MCV_EINVAL, "%s: malformed header", path);
      break;
    }
    b.resize(b.size() * 4);
  }
  const Var *rv[12] = {};
  for (int k = 0; k < 12 && !rc; ++k) {
    for (const Var &v : h.vars)
      if (v.name == kRecordNames[k]) rv[k] = &v;
    if (!rv[k] || !rv[k]->record)
      rc = fail(MCV_EINVAL, "%s has no record variable %s", path, kRecordNames[k]);
    else if (rv[k]->type != (k == 0 ? (uint32_t)kNcInt : (uint32_t)kNcDouble))
      rc = fail(MCV_EINVAL, "%s: variable %s has an unexpected type", path, kRecordNames[k]);
    else if (k > 0 && npoint > 0 &&
             (long long)(rv[k]->vsize / 8) != (k <= 8 ? npoint : ncell))
      rc = fail(MCV_EINVAL, "%s: variable %s holds %lld values per record, this solve %lld",
                path, kRecordNames[k], (long long)(rv[k]->vsize / 8), k <= 8 ? npoint : ncell);
  }
  if (!rc) {
    const uint64_t rec0 = (uint64_t)(time_slot - 1) * h.recsize;
    // output_mod.F90:201: the time variable holds the slot index
    uint8_t t[4] = {(uint8_t)(time_slot >> 24), (uint8_t)(time_slot >> 16),
                    (uint8_t)(time_slot >> 8), (uint8_t)time_slot};
    rc = write_at(f, rv[0]->begin + rec0, t, 4, path);
    for (int k = 1; k < 12 && !rc; ++k)
      rc = write_doubles(f, rv[k]->begin + rec0, src[k - 1], rv[k]->vsize / 8, path);
    if (!rc && (uint32_t)time_slot > h.numrecs) {
      const uint8_t nr[4] = {(uint8_t)(time_slot >> 24), (uint8_t)(time_slot >> 16),
                             (uint8_t)(time_slot >> 8), (uint8_t)time_slot};
      rc = write_at(f, 4, nr, 4, path);
    }
  }
  if (std::fclose(f) != 0 && !rc) rc = fail(MCV_EINVAL, "closing %s failed", path);
  return rc;
}

}  // namespace mcv

