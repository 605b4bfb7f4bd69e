// gfsph_output.hip -- raster output of the propagation results.
//
// NetCDF classic files (the 64-bit-offset variant, "CDF\x02") written by
// hand: the format is a big-endian header of dimensions, attributes and
// variables followed by each variable's values, padded to 4 bytes.  Rasters
// are on the terrain grid: dimensions x (nx) and y (ny), coordinate variables
// x, y (node coordinates xmin + i*dx) and one double variable (y, x) per
// raster, with CF attributes.  This takes the place of the reference's
// hazard-map text outputs (the QGIS and GiD result files written after the
// Monte Carlo loop, "6 SPH_SW_2019.f90":21200-21560) for the statistics the
// batched runs produce.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gfsph.h"

namespace {

enum : uint32_t { NC_CHAR = 2, NC_DOUBLE = 6, NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };

struct Attr {
  std::string name;
  bool is_text;
  std::string text;
  double value;
};

struct Var {
  std::string name;
  std::vector<uint32_t> dims;
  std::vector<Attr> attrs;
  const double *data;
  uint64_t count;
  uint64_t begin = 0;
};

class Buf {
 public:
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void f64(double d) {
    uint64_t v;
    memcpy(&v, &d, 8);
    u64(v);
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void attrs(const std::vector<Attr> &a) {
    if (a.empty()) {
      u32(0);
      u32(0);
      return;
    }
    u32(NC_ATTRIBUTE);
    u32((uint32_t)a.size());
    for (const Attr &t : a) {
      name(t.name);
      if (t.is_text) {
        u32(NC_CHAR);
        u32((uint32_t)t.text.size());
        b.insert(b.end(), t.text.begin(), t.text.end());
        pad();
      } else {
        u32(NC_DOUBLE);
        u32(1);
        f64(t.value);
      }
    }
  }
  std::vector<uint8_t> b;
};

bool valid_name(const char *s) {
  if (!s || !*s) return false;
  const char c0 = s[0];
  if (!((c0 >= 'A' && c0 <= 'Z') || (c0 >= 'a' && c0 <= 'z'))) return false;
  for (const char *p = s; *p; p++) {
    const char c = *p;
    if (!((c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_')) return false;
  }
  return true;
}

// header of the file with the variables' offsets as given
void header(Buf &h, uint32_t nx, uint32_t ny, const std::vector<Attr> &gatts, const std::vector<Var> &vars) {
  const char magic[4] = {'C', 'D', 'F', 2};
  h.b.insert(h.b.end(), magic, magic + 4);
  h.u32(0);  // numrecs: no record variables
  h.u32(NC_DIMENSION);
  h.u32(2);
  h.name("x");
  h.u32(nx);
  h.name("y");
  h.u32(ny);
  h.attrs(gatts);
  h.u32(NC_VARIABLE);
  h.u32((uint32_t)vars.size());
  for (const Var &v : vars) {
    h.name(v.name);
    h.u32((uint32_t)v.dims.size());
    for (uint32_t d : v.dims) h.u32(d);
    h.attrs(v.attrs);
    h.u32(NC_DOUBLE);
    const uint64_t vs = v.count * 8;
    h.u32(vs > 0xfffffffcull ? 0xffffffffu : (uint32_t)vs);
    h.u64(v.begin);
  }
}

}  // namespace

extern "C" {

int gfsph_write_netcdf(const char *path, int32_t nx, int32_t ny, double xmin, double ymin, double dx, double dy,
                       int32_t nvar, const char *const *names, const char *const *units,
                       const char *const *long_names, const double *const *data, double fill, const char *title) {
  if (!path || nx < 1 || ny < 1 || nvar < 0 || (nvar > 0 && (!names || !data))) return GFSPH_ERR_ARG;
  std::vector<double> xs(nx), ys(ny);
  for (int i = 0; i < nx; i++) xs[i] = xmin + (double)i * dx;
  for (int j = 0; j < ny; j++) ys[j] = ymin + (double)j * dy;
  std::vector<Attr> gatts = {{"Conventions", true, "CF-1.8", 0.0},
                             {"title", true, title && *title ? title : "Geoflow-SPH propagation results", 0.0},
                             {"source", true, "gfsph (MI355X depth-integrated SPH propagation)", 0.0}};
  std::vector<Var> vars;
  vars.push_back({"x",
                  {0},
                  {{"standard_name", true, "projection_x_coordinate", 0.0},
                   {"long_name", true, "x coordinate of the terrain grid node", 0.0},
                   {"units", true, "m", 0.0}},
                  xs.data(),
                  (uint64_t)nx});
  vars.push_back({"y",
                  {1},
                  {{"standard_name", true, "projection_y_coordinate", 0.0},
                   {"long_name", true, "y coordinate of the terrain grid node", 0.0},
                   {"units", true, "m", 0.0}},
                  ys.data(),
                  (uint64_t)ny});
  for (int k = 0; k < nvar; k++) {
    if (!valid_name(names[k]) || !strcmp(names[k], "x") || !strcmp(names[k], "y") || !data[k]) return GFSPH_ERR_ARG;
    for (int q = 0; q < k; q++)
      if (!strcmp(names[q], names[k])) return GFSPH_ERR_ARG;
    Var v{names[k], {1, 0}, {}, data[k], (uint64_t)nx * (uint64_t)ny};
    if (long_names && long_names[k] && *long_names[k]) v.attrs.push_back({"long_name", true, long_names[k], 0.0});
    if (units && units[k] && *units[k]) v.attrs.push_back({"units", true, units[k], 0.0});
    if (!std::isnan(fill)) v.attrs.push_back({"_FillValue", false, "", fill});
    vars.push_back(v);
  }
  // offsets: the header's size does not depend on them (8-byte fields)
  Buf h;
  header(h, (uint32_t)nx, (uint32_t)ny, gatts, vars);
  uint64_t off = h.b.size();
  for (Var &v : vars) {
    v.begin = off;
    off += (v.count * 8 + 3) / 4 * 4;
  }
  Buf h2;
  header(h2, (uint32_t)nx, (uint32_t)ny, gatts, vars);
  FILE *f = fopen(path, "wb");
  if (!f) return GFSPH_ERR_ARG;
  bool ok = fwrite(h2.b.data(), 1, h2.b.size(), f) == h2.b.size();
  std::vector<uint8_t> chunk;
  for (const Var &v : vars) {
    const uint64_t chunk_vals = 1 << 16;
    for (uint64_t a = 0; a < v.count && ok; a += chunk_vals) {
      const uint64_t e = std::min<uint64_t>(v.count, a + chunk_vals);
      chunk.resize((e - a) * 8);
      for (uint64_t i = a; i < e; i++) {
        uint64_t u;
        memcpy(&u, &v.data[i], 8);
        for (int s = 0; s < 8; s++) chunk[(i - a) * 8 + s] = (uint8_t)(u >> (56 - 8 * s));
      }
      ok = fwrite(chunk.data(), 1, chunk.size(), f) == chunk.size();
    }
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? GFSPH_OK : GFSPH_ERR_ARG;
}

}  // extern "C"
