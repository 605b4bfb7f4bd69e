// gfsph_inputs.hip -- the reference's input files read on the host and the
// terrain and release built on the GPU.
//
//   <name>.MASTER.dat  Driver_Input run control   ("8 SPH_FEM_Driver_2019.f90":229-470)
//   <name>.dat         Input_SW switches, const(:) ("6 SPH_SW_2019.f90":812-1570)
//   <topo>.top         Topo_main terrain           ("2 SPH_FEM_topo_2019.f90":208-620)
//   <name>.pts         Read_Source_plus_New release (6:4999-5272), or
//   restart.dat        Get_Restart_Info_SW particles (6:3268-3340)
//
// The files are read record by record with Fortran list-directed rules
// (values may continue over records, the rest of the last record is
// dropped, '/' ends a read early, r*v repeats).  Scattered terrain points are
// gridded by Get_DTM's inverse-distance weights (2:1155-1449); nodal slopes
// and curvatures come from Get_Topol's lumped-mass finite-element projection
// on the two-triangle-per-cell mesh (2:583-609, 2:1658-1939); release points
// are gridded by Read_Source_plus_New.  Every accumulation that the
// reference performs as a scatter in input order is a gather here that
// visits its contributions in that same order, so sums round identically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "gfsph.h"

namespace {

// ---------------------------------------------------------------------------
// Fortran list-directed reading
// ---------------------------------------------------------------------------
class FReader {
 public:
  bool open(const std::string &path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string l;
    while (std::getline(f, l)) {
      if (!l.empty() && l.back() == '\r') l.pop_back();
      lines_.push_back(l);
    }
    path_ = path;
    return true;
  }
  // (a60) record
  bool text(std::string *out = nullptr) {
    if (next_ >= lines_.size()) return fail("end of file");
    if (out) *out = lines_[next_];
    next_++;
    return true;
  }
  // list-directed read of n items as strings (fewer if '/' ends the read)
  bool items(size_t n, std::vector<std::string> &tok) {
    tok.clear();
    bool ended = false;
    while (tok.size() < n && !ended) {
      if (next_ >= lines_.size()) return fail("end of file");
      const std::string &l = lines_[next_++];
      size_t k = 0;
      while (k < l.size() && tok.size() < n) {
        while (k < l.size() && (l[k] == ' ' || l[k] == '\t' || l[k] == ',')) k++;
        if (k >= l.size()) break;
        if (l[k] == '/') {
          ended = true;
          break;
        }
        size_t e = k;
        if (l[k] == '\'' || l[k] == '"') {
          const char q = l[k];
          e = l.find(q, k + 1);
          if (e == std::string::npos) e = l.size();
          tok.push_back(l.substr(k + 1, e - k - 1));
          k = e + 1;
          continue;
        }
        while (e < l.size() && l[e] != ' ' && l[e] != '\t' && l[e] != ',' && l[e] != '/') e++;
        std::string t = l.substr(k, e - k);
        const size_t star = t.find('*');
        if (star != std::string::npos && star > 0 &&
            std::all_of(t.begin(), t.begin() + star, [](char ch) { return ch >= '0' && ch <= '9'; })) {
          const long r = atol(t.substr(0, star).c_str());
          const std::string v = t.substr(star + 1);
          for (long q = 0; q < r && tok.size() < n; q++) tok.push_back(v);
        } else {
          tok.push_back(t);
        }
        k = e;
      }
    }
    return true;
  }
  bool ints(std::vector<long> &v, size_t n) {
    std::vector<std::string> t;
    if (!items(n, t)) return false;
    v.assign(n, 0);
    for (size_t k = 0; k < t.size(); k++) {
      char *end = nullptr;
      v[k] = strtol(t[k].c_str(), &end, 10);
      if (end == t[k].c_str() || *end) return fail("integer expected, found '" + t[k] + "'");
    }
    return true;
  }
  bool reals(std::vector<double> &v, size_t n) {
    std::vector<std::string> t;
    if (!items(n, t)) return false;
    v.assign(n, 0.0);
    for (size_t k = 0; k < t.size(); k++)
      if (!real_of(t[k], v[k])) return fail("number expected, found '" + t[k] + "'");
    return true;
  }
  // REAL(4) items: each decimal string rounded once to single precision, as
  // a list-directed read into a REAL(4) variable does
  bool reals4(std::vector<double> &v, size_t n) {
    std::vector<std::string> t;
    if (!items(n, t)) return false;
    v.assign(n, 0.0);
    for (size_t k = 0; k < t.size(); k++) {
      std::string s = fix_exp(t[k]);
      char *end = nullptr;
      const float f = strtof(s.c_str(), &end);
      if (end == s.c_str() || *end) return fail("number expected, found '" + t[k] + "'");
      v[k] = (double)f;
    }
    return true;
  }
  bool logicals(std::vector<int> &v, size_t n) {
    std::vector<std::string> t;
    if (!items(n, t)) return false;
    v.assign(n, 0);
    for (size_t k = 0; k < t.size(); k++) {
      std::string s = t[k];
      size_t a = s[0] == '.' ? 1 : 0;
      if (a >= s.size()) return fail("logical expected");
      const char ch = s[a];
      if (ch == 'T' || ch == 't') v[k] = 1;
      else if (ch == 'F' || ch == 'f') v[k] = 0;
      else return fail("logical expected, found '" + s + "'");
    }
    return true;
  }
  static bool real_of(const std::string &tk, double &out) {
    std::string s = fix_exp(tk);
    char *end = nullptr;
    out = strtod(s.c_str(), &end);
    return end != s.c_str() && !*end;
  }
  const std::string &error() const { return err_; }
  size_t record() const { return next_; }

 private:
  static std::string fix_exp(std::string s) {
    for (char &ch : s)
      if (ch == 'd' || ch == 'D' || ch == 'q' || ch == 'Q') ch = 'e';
    return s;
  }
  bool fail(const std::string &m) {
    err_ = path_ + " record " + std::to_string(next_) + ": " + m;
    return false;
  }
  std::vector<std::string> lines_;
  size_t next_ = 0;
  std::string path_, err_;
};

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// ---------------------------------------------------------------------------
// GPU: ordered window gathers
// ---------------------------------------------------------------------------
#define INP_BLOCK 256
#define MAX_WINDOW 64

// Point indices binned by grid cell: points of bin b are pts[start[b] ..
// start[b+1]), in input order.  Bins are (nbx x nby), row-major.
struct Bins {
  const int *start, *pts;
  int nbx, nby;
};

// Calls f(p) for every point of the bins [bx0, bx1] x [by0, by1] (clamped
// to the bin grid) in increasing point index: a merge of the bins' sorted
// lists.
template <typename F>
__device__ void ordered_window(const Bins &B, int bx0, int bx1, int by0, int by1, F &&f) {
  int cur[MAX_WINDOW], end[MAX_WINDOW];
  int nb = 0;
  bx0 = max(bx0, 0);
  by0 = max(by0, 0);
  bx1 = min(bx1, B.nbx - 1);
  by1 = min(by1, B.nby - 1);
  for (int by = by0; by <= by1; by++)
    for (int bx = bx0; bx <= bx1; bx++) {
      const int b = by * B.nbx + bx;
      if (nb < MAX_WINDOW) {
        cur[nb] = B.start[b];
        end[nb] = B.start[b + 1];
        nb++;
      }
    }
  for (;;) {
    int best = -1, bp = 0x7fffffff;
    for (int k = 0; k < nb; k++)
      if (cur[k] < end[k]) {
        const int p = B.pts[cur[k]];
        if (p < bp) {
          bp = p;
          best = k;
        }
      }
    if (best < 0) break;
    f(bp);
    cur[best]++;
  }
}

// Get_DTM (2:1363-1404): node (ix, iy) of the output grid gathers every DTM
// point whose cell (ipoigx, ipoigy) lies in [ix-2, ix+1] x [iy-2, iy+1]
// (1-based; the reference scatters each point to [ipoigx-1, ipoigx+2] x
// [ipoigy-1, ipoigy+2]) with weight 1 / (max(1e-6, d) / distg).
__global__ void k_dtm_idw(int nx, int ny, double xmin, double ymin, double dxg, double dyg, double distg,
                          const double *px, const double *py, const double *pz, Bins B, double zmaxg, double *Z,
                          int *in_model) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nx * ny) return;
  const int ix = node % nx, iy = node / nx;  // 0-based: ipoigx - 1, ipoigy - 1
  const double xg = xmin + (double)ix * dxg, yg = ymin + (double)iy * dyg;
  double a1 = 0.0, a2 = 0.0;
  // bins are 1-based cell indices ipoigx stored at bx = ipoigx - 1
  ordered_window(B, ix - 2, ix + 1, iy - 2, iy + 1, [&](int p) {
    double dx2 = px[p] - xg;
    dx2 = dx2 * dx2;
    double dy2 = py[p] - yg;
    dy2 = dy2 * dy2;
    double d = fmax(1.e-6, sqrt(dx2 + dy2));
    d = d / distg;
    a1 = a1 + pz[p] / d;
    a2 = a2 + 1.00 / d;
  });
  if (a2 >= 1.e-6) {
    Z[node] = a1 / a2;
    in_model[node] = 1;
  } else {
    Z[node] = zmaxg;
    in_model[node] = 0;
  }
}

// ---------------------------------------------------------------------------
// GPU: Get_Topol (2:1723-1939) on the mesh of 2:583-609.  Cell c (0-based,
// row-major over (nx-1) x (ny-1)) with lower-left node p holds elements 2c
// (p, p+1, p+1+nx) and 2c+1 (p, p+1+nx, p+nx).  A node belongs to at most
// six elements; in increasing element order they are the first and second
// of the cell below-left, the second of the cell below, the first of the
// cell to the left and both of its own cell, which is the order in which
// the reference's element loops add into it.
// ---------------------------------------------------------------------------
struct Mesh {
  int nx, ny;
  double xmin, ymin, dx, dy;
};

__device__ __forceinline__ void elem_nodes(const Mesh &M, int e, int (&n)[3]) {
  const int c = e >> 1, cx = c % (M.nx - 1), cy = c / (M.nx - 1);
  const int p = cy * M.nx + cx;
  n[0] = p;
  if ((e & 1) == 0) {
    n[1] = p + 1;
    n[2] = p + 1 + M.nx;
  } else {
    n[1] = p + 1 + M.nx;
    n[2] = p + M.nx;
  }
}

// getgeom2D (2:1658-1720): shape-function derivatives g[0..2] (x), g[3..5]
// (y) and the Jacobian g[6] of element e
__device__ __forceinline__ void elem_geom(const Mesh &M, int e, const int (&n)[3], double (&g)[7]) {
  double x[3], y[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    x[k] = M.xmin + (double)(n[k] % M.nx) * M.dx;
    y[k] = M.ymin + (double)(n[k] / M.nx) * M.dy;
  }
  const double x21 = x[1] - x[0], x31 = x[2] - x[0], y21 = y[1] - y[0], y31 = y[2] - y[0];
  const double rj = x21 * y31 - x31 * y21;
  const double rj1 = 1. / rj;
  const double xix = y31 * rj1, xiy = -(x31 * rj1), etx = -(y21 * rj1), ety = x21 * rj1;
  const double nxi[3] = {-1.0, 1.0, 0.0}, net[3] = {-1.0, 0.0, 1.0};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    g[k] = xix * nxi[k] + etx * net[k];
    g[k + 3] = xiy * nxi[k] + ety * net[k];
  }
  g[6] = rj;
}

// elements of node `node` in increasing order; returns their count and the
// node's local position (0..2) in each
__device__ __forceinline__ int node_elems(const Mesh &M, int node, int (&es)[6], int (&loc)[6]) {
  const int ix = node % M.nx, iy = node / M.nx;
  const int ncx = M.nx - 1;
  int k = 0;
  if (ix > 0 && iy > 0) {  // cell below-left: node is p+1+nx
    const int c = (iy - 1) * ncx + (ix - 1);
    es[k] = 2 * c;
    loc[k++] = 2;
    es[k] = 2 * c + 1;
    loc[k++] = 1;
  }
  if (ix < ncx && iy > 0) {  // cell below: node is p+nx
    const int c = (iy - 1) * ncx + ix;
    es[k] = 2 * c + 1;
    loc[k++] = 2;
  }
  if (ix > 0 && iy < M.ny - 1) {  // cell to the left: node is p+1
    const int c = iy * ncx + (ix - 1);
    es[k] = 2 * c;
    loc[k++] = 1;
  }
  if (ix < ncx && iy < M.ny - 1) {  // own cell: node is p
    const int c = iy * ncx + ix;
    es[k] = 2 * c;
    loc[k++] = 0;
    es[k] = 2 * c + 1;
    loc[k++] = 0;
  }
  return k;
}

// A1 lumped mass (2:1767-1772)
__global__ void k_topo_mass(Mesh M, double *mml) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= M.nx * M.ny) return;
  int es[6], loc[6];
  const int ne = node_elems(M, node, es, loc);
  double m = 0.0;
  for (int k = 0; k < ne; k++) {
    int n[3];
    double g[7];
    elem_nodes(M, es[k], n);
    elem_geom(M, es[k], n, g);
    m = m + g[6] / 6.0;
  }
  mml[node] = m;
}

// element derivatives of nodal field f (stride fs, offset fo): sum over the
// element's nodes of f * g[k] and f * g[k+3], accumulated from 0 as the
// reference's loops do
__device__ __forceinline__ void elem_grad(const double *f, int fs, int fo, const int (&n)[3], const double (&g)[7],
                                          double &dx, double &dy) {
  dx = 0.0;
  dy = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double v = f[(long)n[k] * fs + fo];
    dx = dx + v * g[k];
    dy = dy + v * g[k + 3];
  }
}

// Right-hand sides of the projections.  stage 0: element gradient of Z
// (A2-A3, 2:1776-1800) -> rhs[2] (x, y); stage 1: element second derivatives
// of (Zx, Zy) (2:1841-1869) -> rhs[3] (xx, (xy+yx)/2, yy).  T is [npoig][4]
// (Z, Zx, Zy, spare) for stage 0 input and the nodal slopes for stage 1.
__global__ void k_topo_rhs(Mesh M, int stage, const double *T, double *rhs) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= M.nx * M.ny) return;
  int es[6], loc[6];
  const int ne = node_elems(M, node, es, loc);
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  for (int k = 0; k < ne; k++) {
    int n[3];
    double g[7];
    elem_nodes(M, es[k], n);
    elem_geom(M, es[k], n, g);
    const double coe = g[6] / 6.;
    if (stage == 0) {
      double dx, dy;
      elem_grad(T, 4, 0, n, g, dx, dy);
      r0 = r0 + dx * coe;
      r1 = r1 + dy * coe;
    } else {
      double xx, xy, yx, yy;
      elem_grad(T, 4, 1, n, g, xx, xy);
      elem_grad(T, 4, 2, n, g, yx, yy);
      const double t5 = (xy + yx) / 2.;
      r0 = r0 + xx * coe;
      r1 = r1 + t5 * coe;
      r2 = r2 + yy * coe;
    }
  }
  rhs[3 * (long)node + 0] = r0;
  rhs[3 * (long)node + 1] = r1;
  rhs[3 * (long)node + 2] = r2;
}

// First estimate v = rhs / m_l (A4 / B4) into v[npoig][3]
__global__ void k_topo_first(int npoig, int nc, const double *rhs, const double *mml, double *v) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= npoig) return;
  const double cm = mml[node];
  for (int d = 0; d < nc; d++) v[3 * (long)node + d] = rhs[3 * (long)node + d] / cm;
}

// One Jacobi iteration v' = v + (rhs - M v) / m_l (A5 / B5, 2:1806-1834,
// 2:1880-1904) with the consistent mass (coe*(1 + [i == j])) gathered per
// node over its elements and, within an element, over its nodes in order.
__global__ void k_topo_iter(Mesh M, int nc, const double *rhs, const double *mml, const double *v, double *vo) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= M.nx * M.ny) return;
  int es[6], loc[6];
  const int ne = node_elems(M, node, es, loc);
  double h[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < ne; k++) {
    int n[3];
    double g[7];
    elem_nodes(M, es[k], n);
    elem_geom(M, es[k], n, g);
    const double coe = g[6] / 24.;
    for (int j = 0; j < 3; j++) {
      const double cm = (j == loc[k]) ? 2.0 : 1.0;
      for (int d = 0; d < nc; d++) h[d] = h[d] + cm * coe * v[3 * (long)n[j] + d];
    }
  }
  const double cm = mml[node];
  for (int d = 0; d < nc; d++) vo[3 * (long)node + d] = v[3 * (long)node + d] + (rhs[3 * (long)node + d] - h[d]) / cm;
}

// Final topol(1:9) (2:1912-1930) into top9[npoig][9]; slopes in s[][3]
// (Zx, Zy, -), curvatures in c[][3]
__global__ void k_topo_final(int npoig, const double *Z, const double *s, const double *c, double *top9,
                             double *T4) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= npoig) return;
  const double zx = s[3 * (long)node], zy = s[3 * (long)node + 1];
  const double zxx = c[3 * (long)node], zxy = c[3 * (long)node + 1], zyy = c[3 * (long)node + 2];
  const double rt = sqrt(1. + zx * zx + zy * zy);
  double *t = top9 + 9 * (long)node;
  t[0] = Z[node];
  t[1] = zx;
  t[2] = zy;
  t[3] = 1. + zx * zx;
  t[4] = zx * zy;
  t[5] = 1. + zy * zy;
  t[6] = zxx / rt;
  t[7] = zxy / rt;
  t[8] = zyy / rt;
  (void)T4;
}

// Z and the slopes packed for the second-derivative stage: T[][4] = (Z, Zx, Zy, 0)
__global__ void k_topo_pack(int npoig, const double *Z, const double *s, double *T) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= npoig) return;
  T[4 * (long)node + 0] = Z[node];
  T[4 * (long)node + 1] = s ? s[3 * (long)node] : 0.0;
  T[4 * (long)node + 2] = s ? s[3 * (long)node + 1] : 0.0;
  T[4 * (long)node + 3] = 0.0;
}

// ---------------------------------------------------------------------------
// GPU: Read_Source_plus_New gridding (6:5166-5212).  Node (inx, iny) gathers
// the release points whose cell (ipoi2Dx, ipoi2Dy) satisfies
// ipoi2Dx - nd_x + 1 <= inx <= ipoi2Dx + nd_x + 1 and
// ipoi2Dy - nd_y <= iny <= ipoi2Dy + nd_y.
// ---------------------------------------------------------------------------
__global__ void k_src_idw(int nx, int ny, double xmin, double ymin, double dxg, double dyg, double dist2D, int ndx,
                          int ndy, const double *px, const double *py, const double *ph, Bins B, double *aux) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nx * ny) return;
  const int ix = node % nx, iy = node / nx;  // 0-based inx - 1, iny - 1
  const double xx0 = xmin + (double)ix * dxg, yy0 = ymin + (double)iy * dyg;
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  ordered_window(B, ix - ndx - 1, ix + ndx - 1, iy - ndy, iy + ndy, [&](int p) {
    double dx2 = px[p] - xx0;
    dx2 = dx2 * dx2;
    double dy2 = py[p] - yy0;
    dy2 = dy2 * dy2;
    double d = fmax(1.e-6, sqrt(dx2 + dy2));
    d = d / dist2D;
    a1 = a1 + 1;
    a2 = a2 + 1.0 / d;
    a3 = a3 + ph[p] / d;
  });
  aux[4 * (long)node + 0] = a1;
  aux[4 * (long)node + 1] = a2;
  aux[4 * (long)node + 2] = a3;
}

// aux2D(4): number of the node's quads whose four corners carry weight
// (6:5195-5212)
__global__ void k_src_area(int nx, int ny, double *aux) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nx * ny) return;
  const int ix = node % nx, iy = node / nx;
  int cnt = 0;
  for (int qy = iy - 1; qy <= iy; qy++)
    for (int qx = ix - 1; qx <= ix; qx++) {
      if (qx < 0 || qy < 0 || qx >= nx - 1 || qy >= ny - 1) continue;
      const int n1 = qy * nx + qx, n2 = n1 + 1, n3 = n1 + nx, n4 = n3 + 1;
      if (aux[4 * (long)n1 + 1] >= 1.e-6 && aux[4 * (long)n2 + 1] >= 1.e-6 && aux[4 * (long)n3 + 1] >= 1.e-6 &&
          aux[4 * (long)n4 + 1] >= 1.e-6)
        cnt++;
    }
  aux[4 * (long)node + 3] = (double)cnt;
}

}  // namespace

// ---------------------------------------------------------------------------
// the inputs object
// ---------------------------------------------------------------------------
struct TopoRaw {
  int ictop = 0, fdiff = 0;
  int topo_props = 0;               // ic_topo_props
  std::vector<gfsph_zone> zones;    // ic_topo_props = 10
  std::vector<double> epts;         // <topo>.eros.pts: [n][3] x, y, depth
  int eros_type = 0;                // type_Eros
  double eros_dxs = 0;              // dx_s
  std::vector<double> gen;          // generated terrains (ictop 1-7, -7): their parameters
  double gx0 = 0, gx1 = 0, gy0 = 0, gy1 = 0;  // Xming, Xmaxg, Yming, Ymaxg read
  std::vector<double> px, py, pz;  // DTM points (ictop 10-16) or grid nodes (ictop 0)
  double deltx = 0;
  double xw0 = 0, xw1 = 0, yw0 = 0, yw1 = 0;
  double zmax = 0;
  std::vector<int> start, pts;  // points binned by cell (ipoigx, ipoigy), in input order
  int nbx = 0, nby = 0;
};

struct SrcRaw {  // Read_Source_plus_New after reading: shifted points and the release grid
  std::vector<double> sx, sy, sh;
  double xmin = 0, ymin = 0, d2x = 0, d2y = 0, dist2D = 0, delt2D = 0, facthsml = 0;
  int nx = 0, ny = 0, ndx = 0, ndy = 0;
  std::vector<int> start, pts;
  int nbx = 0, nby = 0;
};

// a sampled case variable (probab_law, Param_prob_law or the histogram of
// GetMCarlo_Cont_values / GetMCarlo_Dscrt_values, 8:1282-1439)
struct Sampler {
  int law = 0;                    // 0 constant, 1 triangular, 2 normal, 3 log-normal, 100 histogram
  double prm[10] = {};            // Param_prob_law(icase, 1:10)
  std::vector<double> xh, Fh;     // histogram nodes and their accumulated, normalised frequencies
};

// case table of MASTER.dat (Driver_Input, "8 SPH_FEM_Driver_2019.f90":475-1236)
struct CaseTable {
  int mode = 0;  // ic_cases_mat as read: 1 lines, 2 full factorial, 3 sampled rows of it, 4 / 31 / 32
                 // sampled values, 20 / 21 FOSM (first-order second-moment) sets
  int ncases = 0, ninfo = 0;
  std::vector<std::vector<double>> C;        // Ccases_mat rows (10 columns; unread ones 0)
  std::vector<std::vector<double>> list;     // List_Vals rows of statistic-list variables (law >= 100)
  double vstop = 0, timestop = 0, dtimestop = 0;
  std::vector<std::vector<double>> OC;       // OC_mat: [nproblems][3 + ncases + 2]
  // sampled sets: the problems drawn (n_MCarlo_problems), each variable's law
  long nmc = 0;
  std::vector<Sampler> smp;
  std::vector<std::vector<double>> full;     // mode 3: the full factorial table the rows are drawn from
  uint64_t seed = 1;
};

// The uniform numbers of the sampled sets.  The reference draws them with
// Fortran's random_number after RANDOM_SEED(PUT = System_clock count)
// (8:1306-1312), so its sets are not reproducible; here every variable k has
// its own stream from the set's seed: SplitMix64 from state seed +
// 0x9E3779B97F4A7C15 (k + 1), each output u giving ((u >> 11) + 0.5) 2^-53 in
// (0, 1).  mode 3 draws its row indices from stream k = 0.
struct McStream {
  uint64_t s;
  McStream(uint64_t seed, int k) : s(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1)) {}
  double next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return ((double)(z >> 11) + 0.5) * 0x1p-53;
  }
};

// One value of a sampled variable from the uniform y (8:1297-1366,
// 8:1391-1436): the inverse of the law's distribution function.  Normal and
// log-normal use the approximation Phi(t) ~ (1 + sign(t) sqrt(1 - exp(-5 t^2 / 8)))
// / 2 the reference inverts (with y kept in [1e-5, 0.99999]); a histogram's
// distribution is linear between its nodes (y on a node gives the node).
static double mc_value(const Sampler &S, double y) {
  const double *P = S.prm;
  if (S.law == 0) return P[0];
  if (S.law == 1) {
    const double a = P[0], b = P[1], c = P[2];
    const double ba = b - a, ca = c - a, cb = c - b;
    const double baca = ba * ca, cbca = cb * ca, yb = ba / ca;
    return y <= yb ? a + sqrt(baca * y) : c - sqrt(cbca * (1. - y));
  }
  if (S.law == 2 || S.law == 3) {
    const double aa = 5. / 8., aainv = 1 / aa;
    if (y < 0.00001) y = 0.00001;
    if (y > 0.99999) y = 0.99999;
    const double b = log(4 * y * (1. - y));
    const double t = y >= 0.5 ? sqrt(-aainv * b) : -sqrt(-aainv * b);
    if (S.law == 2) return P[0] + P[1] * t;
    const double r = P[1] / P[0];
    const double ls = sqrt(log(1. + r * r));
    const double lm = log(P[0]) - 0.5 * ls * ls;
    return exp(lm + t * ls);
  }
  // histogram
  const size_t np = S.xh.size();
  for (size_t ip = 0; ip + 1 < np; ip++) {
    const double discrim = (S.Fh[ip] - y) * (S.Fh[ip + 1] - y);
    if (discrim < 0.0) {
      const double dd = (y - S.Fh[ip]) / (S.Fh[ip + 1] - S.Fh[ip]);
      return S.xh[ip] + dd * (S.xh[ip + 1] - S.xh[ip]);
    }
    if (S.Fh[ip] == y) return S.xh[ip];
  }
  return S.xh[np - 1];
}

// OC_mat of a sampled set from its seed: modes 4, 31, 32 fill column 3 + k
// with n_MCarlo_problems values of variable k; mode 3 takes n_MCarlo_problems
// rows of the full factorial table at indices nint(y * nproblems), kept in
// 1 .. nproblems (8:623-630).
static void mc_fill(CaseTable &T) {
  const int ncol = 3 + T.ncases + 2;
  if (T.mode == 3) {
    const long np = (long)T.full.size();
    McStream st(T.seed, 0);
    T.OC.clear();
    for (long k = 0; k < T.nmc; k++) {
      long ir = lround(st.next() * (double)np);  // nint: halves away from zero
      if (ir == 0) ir = 1;
      if (ir > np) ir = np;
      T.OC.push_back(T.full[ir - 1]);
    }
    return;
  }
  T.OC.assign(T.nmc, std::vector<double>(ncol, 0.0));
  for (int k = 0; k < T.ncases; k++) {
    McStream st(T.seed, k);
    for (long p = 0; p < T.nmc; p++) T.OC[p][3 + k] = mc_value(T.smp[k], st.next());
  }
}

struct gfsph_inputs {
  std::string err;
  gfsph_params P{};
  gfsph_run run{};
  CaseTable cases;
  gfsph_grid grid{};
  TopoRaw topo;
  SrcRaw src;
  bool built = false;
  std::vector<double> Z, topol;  // [npoig], [npoig][9] (empty when slopes are finite differences)
  std::vector<int32_t> in_model;
  std::vector<int32_t> zone_nodes;  // node-list bits of geometry-101 zones (empty: none)
  std::vector<double> eros_avail;   // topol(15,:) after the terrain zones (empty: none)
  int n = 0;
  std::vector<double> x, v, h, mass, hsml, u;
  std::vector<int32_t> itype;
  int icunk = 0;
};

namespace {

int ifail(gfsph_inputs *in, int code, const std::string &m) {
  in->err = m;
  return code;
}

#define RD(call)                                 \
  do {                                           \
    if (!(call)) return ifail(in, GFSPH_ERR_ARG, R.error()); \
  } while (0)

// One variable's law: probab_law, then Param_prob_law(1:10) for a
// continuous law or the histogram (np_h, nodes x_h, element counts p_h) that
// GetMCarlo_Dscrt_values reads (8:735-765, 8:1391-1406).
int read_law(gfsph_inputs *in, FReader &R, const std::string &path, Sampler &S) {
  std::vector<long> iv;
  std::vector<double> rv;
  RD(R.text());
  RD(R.ints(iv, 1));
  S.law = (int)iv[0];
  if (S.law < 100) {
    RD(R.text());
    RD(R.reals(rv, 10));
    for (int k = 0; k < 10; k++) S.prm[k] = rv[k];
  } else if (S.law == 100) {
    RD(R.text());
    RD(R.ints(iv, 1));
    const long np = iv[0];
    if (np < 2) return ifail(in, GFSPH_ERR_ARG, path + ": a histogram needs at least 2 nodes");
    RD(R.text());
    RD(R.reals(rv, np));
    S.xh = rv;
    RD(R.text());
    RD(R.reals(rv, np - 1));
    S.Fh.assign(np, 0.0);
    for (long ip = 1; ip < np; ip++) S.Fh[ip] = S.Fh[ip - 1] + rv[ip - 1];
    const double tot = S.Fh[np - 1];
    if (!(tot > 0)) return ifail(in, GFSPH_ERR_ARG, path + ": histogram frequencies must add up to > 0");
    for (long ip = 0; ip < np; ip++) S.Fh[ip] = S.Fh[ip] / tot;
  }
  return GFSPH_OK;
}

// the laws the samplers draw from (others leave the values 0 in the reference)
int check_law(gfsph_inputs *in, const std::string &path, const Sampler &S) {
  if (S.law == 1 && !(S.prm[2] > S.prm[0]))
    return ifail(in, GFSPH_ERR_ARG, path + ": a triangular law needs max > min");
  if (S.law == 3 && !(S.prm[0] > 0))
    return ifail(in, GFSPH_ERR_ARG, path + ": a log-normal law needs a positive mean");
  if (!(S.law == 0 || S.law == 1 || S.law == 2 || S.law == 3 || S.law == 100))
    return ifail(in, GFSPH_ERR_UNSUPPORTED,
                 path + ": probability law " + std::to_string(S.law) +
                     " (0 constant, 1 triangular, 2 normal, 3 log-normal, 100 histogram are drawn)");
  return GFSPH_OK;
}

// ic_cases_mat 32 (8:1138-1234): the stop thresholds, ncases_mat, mx_cases,
// ninfo_mat, n_MCarlo_problems and n_GZones (taken as 1), then mx_cases
// records (variable, law, 10 parameters); the records with a law > 0 become
// the case variables in order, continuous laws only (a histogram stops the
// reference).
int read_case_set32(gfsph_inputs *in, FReader &R, const std::string &path) {
  CaseTable &T = in->cases;
  std::vector<long> iv;
  std::vector<double> rv;
  RD(R.text());
  RD(R.reals(rv, 3));
  T.vstop = rv[0];
  T.timestop = rv[1];
  T.dtimestop = rv[2];
  RD(R.text());
  RD(R.ints(iv, 3));
  T.ncases = (int)iv[0];
  const long mx = iv[1];
  T.ninfo = (int)iv[2];
  if (T.ncases < 1 || mx < T.ncases) return ifail(in, GFSPH_ERR_ARG, path + ": 1 <= ncases_mat <= mx_cases expected");
  RD(R.text());
  RD(R.ints(iv, 2));
  T.nmc = iv[0];
  if (T.nmc < 1 || T.nmc > 1000000) return ifail(in, GFSPH_ERR_ARG, path + ": 1 to 10^6 Monte Carlo problems");
  T.C.assign(T.ncases, std::vector<double>(10, 0.0));
  T.smp.assign(T.ncases, Sampler());
  RD(R.text());
  int nc = 0;
  for (long r = 0; r < mx; r++) {
    RD(R.reals(rv, 12));
    if ((long)rv[1] > 0 && nc < T.ncases) {
      T.C[nc][0] = rv[0];
      T.smp[nc].law = (int)rv[1];
      for (int k = 0; k < 10; k++) T.smp[nc].prm[k] = rv[2 + k];
      nc++;
    }
  }
  if (nc < T.ncases) return ifail(in, GFSPH_ERR_ARG, path + ": fewer records with a law than ncases_mat");
  for (int c = 0; c < T.ncases; c++) {
    if (T.smp[c].law >= 100)
      return ifail(in, GFSPH_ERR_ARG, path + ": set 32 draws from continuous laws only (8:1218-1222)");
    if (int rc = check_law(in, path, T.smp[c])) return rc;
  }
  T.list.assign(T.ncases, std::vector<double>());
  mc_fill(T);
  return GFSPH_OK;
}

// Driver_Input subset (8:229-340): SPH only, no GFL, no case tables
// Case tables and the problem matrix OC_mat (8:475-639): mode 1 varies one
// variable at a time over its values (the others at their defaults), mode 2
// takes every combination (variable 1 fastest); values are min + (max -
// min)(i - 1)/(n - 1), on a log scale for law 2 (and, in mode 1, for const 13
// and 27), or the i-th entry of a statistic list (law >= 100, mode 2).
int read_case_table(gfsph_inputs *in, FReader &R, const std::string &path) {
  CaseTable &T = in->cases;
  std::vector<long> iv;
  std::vector<double> rv;
  if (T.mode == 32) return read_case_set32(in, R, path);
  RD(R.text());
  RD(R.ints(iv, 2));
  T.ncases = (int)iv[0];
  T.ninfo = (int)iv[1];
  if (T.ncases < 1 || T.ninfo < 5 || T.ninfo > 10)
    return ifail(in, GFSPH_ERR_ARG, path + ": ncases_mat >= 1 and 5 <= ninfo_mat <= 10 expected");
  RD(R.text());
  T.C.assign(T.ncases, std::vector<double>(10, 0.0));
  for (int c = 0; c < T.ncases; c++) {
    RD(R.reals(rv, T.ninfo));
    for (int k = 0; k < T.ninfo; k++) T.C[c][k] = rv[k];
  }
  T.list.assign(T.ncases, std::vector<double>());
  if (T.mode == 2 || T.mode == 3) {
    bool any = false;
    for (int c = 0; c < T.ncases; c++) any = any || T.C[c][5] >= 100;
    if (any) {
      RD(R.text());
      for (int c = 0; c < T.ncases; c++) {
        if (!(T.C[c][5] >= 100)) continue;
        const long nval = (long)T.C[c][2];
        RD(R.reals(rv, nval + 1));
        T.list[c] = rv;
        RD(R.reals(rv, nval + 1));  // List_Probs: not used by the deterministic table
      }
    }
  }
  RD(R.text());
  RD(R.reals(rv, 3));
  T.vstop = rv[0];
  T.timestop = rv[1];
  T.dtimestop = rv[2];
  const int ncol = 3 + T.ncases + 2;
  T.OC.clear();
  if (T.mode == 4 || T.mode == 5 || T.mode == 31) {  // sampled values (8:693-857, 8:1053-1136)
    RD(R.text());
    long nz = 1;
    if (T.mode == 31 || T.mode == 5) {
      RD(R.ints(iv, 2));
      nz = iv[1];
    } else {
      RD(R.ints(iv, 1));
    }
    T.nmc = iv[0];
    if (T.nmc < 1 || T.nmc > 1000000) return ifail(in, GFSPH_ERR_ARG, path + ": 1 to 10^6 Monte Carlo problems");
    if (nz != 1)
      return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": n_GZones > 1 (the terrain's geological zones) is not read");
    T.smp.assign(T.ncases, Sampler());
    for (int c = 0; c < T.ncases; c++) {
      int rc = read_law(in, R, path, T.smp[c]);
      if (!rc) rc = check_law(in, path, T.smp[c]);
      if (rc) return rc;
    }
    mc_fill(T);
    return GFSPH_OK;
  }
  if (T.mode == 20 || T.mode == 21) {  // FOSM: the means and the sigmas, 2 ncases + 1 problems (8:974-1051)
    RD(R.text());
    std::vector<double> mu, sg;
    RD(R.reals(mu, T.ncases));
    RD(R.reals(sg, T.ncases));
    T.OC.assign(2 * T.ncases + 1, std::vector<double>(ncol, 0.0));
    for (int c = 0; c < T.ncases; c++) T.OC[0][3 + c] = mu[c];
    // as the reference fills rows 2c and 2c+1: every column the mean of
    // variable c (8:1039-1045), then column c at mean +- sigma/10
    for (int c = 0; c < T.ncases; c++)
      for (int j = 0; j < T.ncases; j++) T.OC[2 * c + 1][3 + j] = T.OC[2 * c + 2][3 + j] = mu[c];
    for (int c = 0; c < T.ncases; c++) {
      T.OC[2 * c + 1][3 + c] = mu[c] + sg[c] / 10.;
      T.OC[2 * c + 2][3 + c] = mu[c] - sg[c] / 10.;
    }
    return GFSPH_OK;
  }
  auto value = [&](int c, long ival, int col_lo, bool logs) {  // min + dval*(ival-1), log scales
    double mn = T.C[c][col_lo], mx = T.C[c][col_lo + 1];
    const long nval = (long)T.C[c][2];
    if (logs) {
      mn = log(mn);
      mx = log(mx);
    }
    const double dv = nval == 1 ? 0.0 : (mx - mn) / (nval - 1);
    double v = mn + dv * (ival - 1);
    if (logs) v = exp(v);
    return v;
  };
  if (T.mode == 1) {
    for (int c = 0; c < T.ncases; c++) {
      const int indx = (int)T.C[c][0];
      const long nval = (long)T.C[c][2];
      for (long ival = 1; ival <= nval; ival++) {
        std::vector<double> row(ncol, 0.0);
        row[0] = c + 1;
        row[1] = (double)ival;
        row[2] = indx;
        const double v = value(c, ival, 3, indx == 13 || indx == 27);
        for (int j = 0; j < T.ncases; j++) row[3 + j] = (j == c) ? v : T.C[j][1];
        T.OC.push_back(row);
      }
    }
  } else {
    long np = 1;
    std::vector<long> stride(T.ncases);
    for (int c = 0; c < T.ncases; c++) {
      const long nval = (long)T.C[c][2];
      if (nval < 1) return ifail(in, GFSPH_ERR_ARG, path + ": every case variable needs at least one value");
      stride[c] = np;
      np *= nval;
    }
    if (np > 1000000) return ifail(in, GFSPH_ERR_ARG, path + ": more than 10^6 problems");
    for (long ip = 0; ip < np; ip++) {
      std::vector<double> row(ncol, 0.0);
      for (int c = 0; c < T.ncases; c++) {
        const long ival = (ip / stride[c]) % (long)T.C[c][2] + 1;
        const int law = (int)T.C[c][5];
        if (law >= 100) {
          if ((long)T.list[c].size() < ival + 1) return ifail(in, GFSPH_ERR_ARG, path + ": statistic list too short");
          row[3 + c] = T.list[c][ival];
        } else {
          row[3 + c] = value(c, ival, 3, law == 2);
          if ((int)T.C[c][0] == 521) row[3 + T.ncases + 1] = value(c, ival, 6, false);
        }
      }
      T.OC.push_back(row);
    }
    if (T.mode == 3) {  // n_MCarlo_problems rows drawn from the table (8:589-600, 8:623-634)
      RD(R.text());
      RD(R.ints(iv, 1));
      T.nmc = iv[0] < np ? iv[0] : np;
      if (T.nmc < 1) return ifail(in, GFSPH_ERR_ARG, path + ": n_MCarlo_problems >= 1 expected");
      T.full.swap(T.OC);
      mc_fill(T);
    }
  }
  return GFSPH_OK;
}

int read_master(gfsph_inputs *in, const std::string &path) {
  FReader R;
  if (!R.open(path)) return GFSPH_OK;  // optional
  std::vector<long> iv;
  std::vector<double> rv;
  RD(R.ints(iv, 1));
  for (long k = 0; k < iv[0]; k++) RD(R.text());
  RD(R.text());
  RD(R.ints(iv, 3));
  const long if_sph = iv[0], if_gfl = iv[1], cases = iv[2];
  if (if_sph != 1 || if_gfl != 0)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": only if_sph = 1, if_gfl = 0 runs are read");
  if (!(cases == 0 || cases == 1 || cases == 2 || cases == 3 || cases == 4 || cases == 5 || cases == 20 ||
        cases == 21 || cases == 31 || cases == 32))
    return ifail(in, GFSPH_ERR_UNSUPPORTED,
                 path + ": ic_cases_mat " + std::to_string(cases) +
                     " (0-5, 20, 21, 31, 32 are read; 6 is the trigger module's FOSM with return periods)");
  in->cases.mode = (int)cases;
  RD(R.text());
  RD(R.ints(iv, 2));
  if (iv[0] != 1) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": SPH_ProblemType must be 1 (shallow water)");
  in->run.t_integ_alg = (int32_t)iv[1];
  RD(R.text());
  RD(R.text());  // problem name
  RD(R.text());
  RD(R.reals(rv, 3));
  in->run.dt = rv[0];
  in->run.time_end = rv[1];
  in->run.maxtimestep = (int32_t)rv[2];
  RD(R.text());
  RD(R.ints(iv, 3));
  in->run.print_step = (int32_t)std::labs(iv[0]);
  in->run.save_step = (int32_t)iv[1];
  in->run.plot_step = (int32_t)iv[2];
  if (iv[0] < 0) {
    RD(R.text());
    RD(R.ints(iv, 1));
  }
  RD(R.text());
  std::vector<std::string> t;
  RD(R.items(2, t));
  if (t.size() < 2 || !FReader::real_of(t[0], in->run.dt_sph))
    return ifail(in, GFSPH_ERR_ARG, path + ": dt_sph, ic_adapt_dt_sph expected");
  in->run.adapt_dt = atoi(t[1].c_str());
  in->run.present = 1;
  // time curves (8:335-357)
  RD(R.text());
  RD(R.ints(iv, 2));
  const long ntc = iv[0];
  for (long k = 0; k < ntc; k++) {
    RD(R.text());
    RD(R.ints(iv, 1));
    const long np = iv[0];
    RD(R.text());
    RD(R.reals(rv, np));
    RD(R.reals(rv, np));
  }
  if (cases) return read_case_table(in, R, path);
  return GFSPH_OK;
}

// Input_SW (6:812-1570) for a single-phase soil release
int generated_release(gfsph_inputs *in, FReader &R, const std::string &path);

int read_dat(gfsph_inputs *in, const std::string &path, std::string &src_name) {
  FReader R;
  if (!R.open(path)) return ifail(in, GFSPH_ERR_ARG, "cannot open " + path);
  std::vector<long> iv;
  std::vector<double> rv;
  std::vector<int> lv;
  gfsph_params &P = in->P;
  RD(R.ints(iv, 1));
  for (long k = 0; k < iv[0]; k++) RD(R.text());
  RD(R.text());
  RD(R.ints(iv, 1));
  if (iv[0] != 0) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": ic_SWAlg must be 0");
  RD(R.text());
  RD(R.ints(iv, 1));
  if (iv[0] != 0) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": histograms (n_hist > 0) are not read");
  RD(R.text());
  RD(R.ints(iv, 1));
  if (iv[0] != 2) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": ndimn must be 2");
  RD(R.text());
  RD(R.ints(iv, 4));
  if (iv[0] != 1 || iv[1] != 0 || iv[2] != 0 || iv[3] != 0)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": only a soil release (1 0 0 0) is read");
  // Get_soil_points (6:3207-3260)
  RD(R.text());
  std::vector<std::string> t;
  RD(R.items(2, t));
  if (t.size() < 2) return ifail(in, GFSPH_ERR_ARG, path + ": icunk_s, h_inf_SW expected");
  in->icunk = atoi(t[0].c_str());
  if (!FReader::real_of(t[1], P.h_inf)) return ifail(in, GFSPH_ERR_ARG, path + ": h_inf_SW");
  const bool gen = in->icunk == 1 || in->icunk == 11 || in->icunk == 12 || in->icunk == 13;
  if (in->icunk != 7 && in->icunk != 200 && !gen)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": release type icunk_s " + std::to_string(in->icunk) +
                                                " (1: cylinder, 7: gridded .pts, 11: rectangle, 12: ellipse, 13:"
                                                " rectangle with a paraboloid depth and 200: restart.dat are read)");
  if (gen) {
    int rc = generated_release(in, R, path);
    if (rc) return rc;
  } else {
    RD(R.text());
    std::string nm;
    RD(R.text(&nm));
    src_name = trim(nm);
  }
  // switches (6:1101-1110)
  RD(R.text());
  RD(R.ints(iv, 4));
  P.pa_sph = (int32_t)iv[0];
  P.nnps = (int32_t)iv[1];
  P.sle = (int32_t)iv[2];
  P.skf = (int32_t)iv[3];
  RD(R.text());
  RD(R.logicals(lv, 4));
  P.summation_density = lv[0];
  P.average_velocity = lv[1];
  if (lv[2]) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": virtual particles are not supported");
  P.nor_density = lv[3];
  // const(:) (6:1130-1150), REAL(4)
  RD(R.text());
  RD(R.reals4(rv, 15));
  for (int k = 1; k <= 15; k++) P.cst[k] = rv[k - 1];
  if (P.cst[4] > 1.e-4) {
    P.cst[21] = 1;
    P.cst[22] = P.cst[4];
  }
  if (P.cst[11] < 0) {
    RD(R.text());
    RD(R.reals4(rv, 15));
    for (int k = 16; k <= 30; k++) P.cst[k] = rv[k - 16];
  }
  if (P.cst[30] < 0) {
    RD(R.text());
    RD(R.reals4(rv, 15));
    for (int k = 31; k <= 45; k++) P.cst[k] = rv[k - 31];
  }
  if (P.cst[45] < 0) {
    RD(R.text());
    RD(R.reals4(rv, 15));
    for (int k = 46; k <= 60; k++) P.cst[k] = rv[k - 46];
  }
  const int nfric = (int)(P.cst[5] + 0.01);
  const bool frictional = (nfric >= 5 && nfric <= 10) || nfric == 26 || nfric == 27;
  if (nfric == 201 || nfric == 202 || nfric == 203)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": nfric 201-203 (water inflow) is not supported");
  if (P.cst[1] < 0) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": inclined gravity (const(1) < 0)");
  if (frictional) {  // ic_K0 and its coefficients (6:1373-1390)
    RD(R.text());
    RD(R.ints(iv, 1));
    P.ic_k0 = (int32_t)iv[0];
    if (P.ic_k0 == 1) {
      RD(R.text());
      RD(R.reals(rv, 2));
      P.k0_active = rv[0];
      P.k0_pasive = rv[1];
    }
  }
  RD(R.text());
  RD(R.ints(iv, 1));
  P.icpwp = (int32_t)iv[0];
  if (P.icpwp == 1 || P.icpwp == 3) {  // Initialize_Pwp (6:1722-1729)
    RD(R.text());
    RD(R.reals(rv, 1));
    P.pwp0rel = rv[0];
  } else if (P.icpwp == 2) {  // Initialize_Pwp, the pore-pressure column (6:1743-1768)
    RD(R.text());
    RD(R.ints(iv, 1));
    P.nuaux = (int32_t)iv[0];
    RD(R.text());
    RD(R.reals(rv, 4));  // BcT1 BcV1 BcT2 BcV2: the types are integers
    if (rv[0] != (double)(long)rv[0] || rv[2] != (double)(long)rv[2])
      return ifail(in, GFSPH_ERR_ARG, path + ": pore-pressure condition types must be integers");
    P.pwp_bc_type[0] = (int32_t)rv[0];
    P.pwp_bc_val[0] = rv[1];
    P.pwp_bc_type[1] = (int32_t)rv[2];
    P.pwp_bc_val[1] = rv[3];
    RD(R.text());
    RD(R.reals(rv, 2));  // pwp0_type (a REAL in the reference) and pwp0rel
    P.pwp0_type = (int32_t)rv[0];
    P.pwp0rel = rv[1];
    if (P.pwp0_type == 4) {
      RD(R.text());
      RD(R.reals(rv, 1));
      P.xirel = rv[0];
    }
  } else if (P.icpwp == 20) {
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": icpwp = 20 (pore pressure with Stefan erosion) is not supported");
  } else if (P.icpwp != 0) {
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": icpwp must be 0, 1, 2 or 3");
  }
  P.integrator = 4;
  P.init_input = 1;
  return GFSPH_OK;
}

// A release the reference generates from a few numbers of the .dat file:
// Read_rect_SW (icunk_s 11, "6 SPH_SW_2019.f90":4426-4497) -- npoinx x
// npoiny points over the rectangle of axis (x0, y0)-(x1, y1) and half-width
// b, depth d0, equal masses 4ab d0 / n, hsml sqrt(4ab/n) facthsml -- and
// Read_ellipsoidal_SW (icunk_s 12, 6:3818-3899) -- the points of an ellipse
// of half-axes a, b turned theta0 about (xp0, yp0), rows of spacing drr
// across the major axis, depth hp0 ((1 - r) + 0.1 r) for r the elliptic
// radius squared, mass the depth times the row's cell -- and
// Read_rect_ellipsoid (icunk_s 13, 6:3903-3979: the rectangle's points with
// the depth d0 (1 - xi^2/a^2)(1 - eta^2/b^2) + sqrt(d0^2/n)) and
// Read_Cylinder_SW (icunk_s 1, 6:3983-4089: a centre point and rings of
// 6k points at radius k rp0/nlinep, depth constant, a hemisphere-like
// sqrt(1 - r^2/rp0^2) or a second depth inside rp1).  At rest, u = 0.
int generated_release(gfsph_inputs *in, FReader &R, const std::string &path) {
  std::vector<long> iv;
  std::vector<double> rv;
  auto add = [&](double x, double y, double h, double m, double hs) {
    in->x.push_back(x);
    in->x.push_back(y);
    in->v.push_back(0.0);
    in->v.push_back(0.0);
    in->h.push_back(h);
    in->mass.push_back(m);
    in->hsml.push_back(hs);
    in->u.push_back(0.0);
    in->itype.push_back(1);
  };
  if (in->icunk == 1) {
    RD(R.text());
    std::vector<std::string> t;
    RD(R.items(5, t));
    double c5[4];
    for (int k = 0; k < 4; k++)
      if (!FReader::real_of(t[k], c5[k])) return ifail(in, GFSPH_ERR_ARG, path + ": xp0, yp0, rp0, hp0");
    const double xp0 = c5[0], yp0 = c5[1], rp0 = c5[2], hp0 = c5[3];
    const int icdens = atoi(t[4].c_str());
    if (icdens < 0 || icdens > 2) return ifail(in, GFSPH_ERR_ARG, path + ": icdens must be 0, 1 or 2");
    double rp1 = 0, hp1 = 0, fp0 = 1.0;
    if (icdens == 2) {
      RD(R.text());
      RD(R.reals(rv, 2));
      rp1 = rv[0];
      hp1 = rv[1];
      fp0 = hp1 / hp0;
    }
    RD(R.text());
    RD(R.items(2, t));
    const long nlinep = atol(t[0].c_str());
    double fh;
    if (nlinep < 0 || !FReader::real_of(t[1], fh)) return ifail(in, GFSPH_ERR_ARG, path + ": nlinep, facthsml");
    const double pi = 4. * atan(1.0);
    const long np = 1 + 3 * nlinep * (nlinep + 1);
    const double mss = pi * rp0 * rp0 * hp0 / (double)np;
    const double drr = nlinep >= 1 ? rp0 / (double)nlinep : rp0;
    add(xp0, yp0, hp0 * fp0, mss * fp0, 2. * drr);
    for (long il = 1; il <= nlinep; il++) {
      const double delth = 2. * pi / (6. * (double)il);
      double rpp = (double)il * drr;
      if ((rpp / rp0) > 1.00) rpp = 0.9999999 * rp0;
      const double f1 = (icdens == 2 && rpp <= rp1) ? hp1 / hp0 : 1.0;
      double hfact = sqrt(1.0 - (rpp / rp0) * (rpp / rp0));
      if (hfact < 0.1) hfact = 0.1;
      for (long ith = 1; ith <= 6 * il; ith++) {
        const double theta = (double)ith * delth;
        const double h = icdens == 0 ? hp0 : (icdens == 1 ? hp0 * hfact : hp0 * f1);
        const double m = icdens == 0 ? mss : (icdens == 1 ? mss * hfact : mss * f1);
        add(xp0 + rpp * cos(theta), yp0 + rpp * sin(theta), h, m, 2. * drr * fh);
      }
    }
  } else if (in->icunk == 11 || in->icunk == 13) {
    RD(R.text());
    RD(R.ints(iv, 2));
    const long npx = iv[0], npy = iv[1];
    if (npx < 2 || npy < 2) return ifail(in, GFSPH_ERR_ARG, path + ": npoinx, npoiny must be >= 2");
    RD(R.text());
    RD(R.reals(rv, 7));
    const double x0 = rv[0], y0 = rv[1], x1 = rv[2], y1 = rv[3], b = rv[4], d0 = rv[5], fh = rv[6];
    const double xc = (x0 + x1) / 2., yc = (y0 + y1) / 2.;
    const double a = sqrt((x1 - xc) * (x1 - xc) + (y1 - yc) * (y1 - yc));
    const double vnxi = (x1 - xc) / a, vneta = (y1 - yc) / a, vvxi = -vneta, vveta = vnxi;
    const double xs = x0 + b * vneta, ys = y0 - b * vnxi;
    const long np = npx * npy;
    const double area = 4 * a * b / (double)np, mss = area * d0, hs = sqrt(area) * fh;
    const double hmin = sqrt(d0 * d0 / (double)np);  // icunk_s 13 outside the rectangle's ellipsoid
    for (long jy = 1; jy <= npy; jy++)
      for (long jx = 1; jx <= npx; jx++) {
        const double x = xs + 2. * a * vnxi * (double)(jx - 1) / (double)(npx - 1) +
                         2. * b * vvxi * (double)(jy - 1) / (double)(npy - 1);
        const double y = ys + 2. * a * vneta * (double)(jx - 1) / (double)(npx - 1) +
                         2. * b * vveta * (double)(jy - 1) / (double)(npy - 1);
        if (in->icunk == 11) {
          add(x, y, d0, mss, hs);
        } else {
          const double xi = (x - xc) * vnxi + (y - yc) * vneta, eta = (x - xc) * vvxi + (y - yc) * vveta;
          double h = hmin;
          if (fabs(xi) <= a && fabs(eta) <= b) h = d0 * (1 - xi / a) * (1 + xi / a) * (1 - eta / b) * (1 + eta / b) + hmin;
          add(x, y, h, h * area, hs);
        }
      }
  } else {
    RD(R.text());
    RD(R.reals(rv, 6));
    const double xp0 = rv[0], yp0 = rv[1], a = rv[2], b = rv[3], th = rv[4], hp0 = rv[5];
    RD(R.text());
    RD(R.reals(rv, 2));
    const double drr = rv[0], fh = rv[1];
    if (!(drr > 0) || !(a > 0) || !(b > 0)) return ifail(in, GFSPH_ERR_ARG, path + ": a, b, drr must be > 0");
    const int nxi = (int)(2 * a / drr + 1);
    if (nxi < 2) return ifail(in, GFSPH_ERR_ARG, path + ": the ellipse is shorter than two rows");
    const double dxi = 2 * a / (nxi - 1);
    for (int ixi = 1; ixi <= nxi; ixi++) {
      const double xi = -a + (ixi - 1) * dxi;
      double bmin = (1 - xi / a) * (1 + xi / a);
      if (bmin > 0.0) bmin = b * sqrt(bmin);
      const int neta = (int)(2 * bmin / drr + 1);
      const double deta = neta > 1 ? 2 * bmin / (neta - 1) : drr;
      const double dArea = dxi * deta;
      for (int ieta = 1; ieta <= neta; ieta++) {
        const double eta = -bmin + (ieta - 1) * deta;
        const double rpp = (xi / a) * (xi / a) + (eta / b) * (eta / b);
        if (rpp <= 1) {
          const double hfact = (1. - rpp) + 0.1 * rpp;
          const double h = hp0 * hfact;
          add(xp0 + xi * cos(th) - eta * sin(th), yp0 + xi * sin(th) + eta * cos(th), h, h * dArea, drr * fh);
        }
      }
    }
  }
  in->n = (int)in->h.size();
  return GFSPH_OK;
}

// Topo_main (2:208-252, 2:330-460) and Get_DTM's reading (2:1181-1227)
int read_top(gfsph_inputs *in, const std::string &path, TopoRaw &T) {
  FReader R;
  if (!R.open(path)) return ifail(in, GFSPH_ERR_ARG, "cannot open " + path);
  std::vector<long> iv;
  std::vector<double> rv;
  RD(R.text());
  RD(R.ints(iv, 1));
  T.ictop = (int)iv[0];
  if (T.ictop >= 100) {
    T.fdiff = 1;
    T.ictop -= 100;
  }
  if (T.ictop == 0) {
    RD(R.text());
    RD(R.ints(iv, 2));
    in->grid.nx = (int32_t)iv[0];
    in->grid.ny = (int32_t)iv[1];
    RD(R.text());
    const long np = (long)in->grid.nx * in->grid.ny;
    if (np <= 0) return ifail(in, GFSPH_ERR_ARG, path + ": npoigx, npoigy");
    T.px.assign(np, 0.0);
    T.py.assign(np, 0.0);
    T.pz.assign(np, 0.0);
    for (long k = 0; k < np; k++) {
      RD(R.reals(rv, 4));
      const long ip = (long)rv[0] - 1;
      if (ip < 0 || ip >= np) return ifail(in, GFSPH_ERR_ARG, path + ": node index out of range");
      T.px[ip] = rv[1];
      T.py[ip] = rv[2];
      T.pz[ip] = rv[3];
    }
  } else if ((T.ictop >= 1 && T.ictop <= 7) || T.ictop == -7) {
    // a grid generated from a formula (2:260-444): the extent and node
    // counts, then the formula's parameters
    RD(R.text());
    RD(R.reals(rv, 4));
    T.gx0 = rv[0];
    T.gx1 = rv[1];
    T.gy0 = rv[2];
    T.gy1 = rv[3];
    RD(R.text());
    RD(R.ints(iv, 2));
    in->grid.nx = (int32_t)iv[0];
    in->grid.ny = (int32_t)iv[1];
    if (in->grid.nx < 2 || in->grid.ny < 2) return ifail(in, GFSPH_ERR_ARG, path + ": npoigx, npoigy");
    RD(R.text());
    if (T.ictop == 4) {  // multilinear in x: nptsZ, then the x's and the Z's
      RD(R.ints(iv, 1));
      const long nz = iv[0];
      if (nz < 2) return ifail(in, GFSPH_ERR_ARG, path + ": nptsZ");
      RD(R.text());
      RD(R.reals(rv, nz));
      T.gen.assign(rv.begin(), rv.begin() + nz);
      RD(R.reals(rv, nz));
      T.gen.insert(T.gen.end(), rv.begin(), rv.begin() + nz);
    } else {
      const int nr = T.ictop == 1 ? 1 : (T.ictop == 3 ? 6 : (T.ictop == 7 || T.ictop == -7 ? 2 : 4));
      RD(R.reals(rv, nr));
      T.gen.assign(rv.begin(), rv.begin() + nr);
    }
  } else if (T.ictop == 10 || T.ictop == 11 || T.ictop == 15 || T.ictop == 16) {
    RD(R.text());
    RD(R.reals(rv, 2));
    const long np = (long)rv[0];
    T.deltx = rv[1];
    if (np <= 0 || !(T.deltx > 0)) return ifail(in, GFSPH_ERR_ARG, path + ": np_dtm, deltx");
    if (T.ictop == 15 || T.ictop == 16) {
      RD(R.text());
      RD(R.reals(rv, 4));
      T.xw0 = rv[0];
      T.xw1 = rv[1];
      T.yw0 = rv[2];
      T.yw1 = rv[3];
    }
    RD(R.text());
    T.px.resize(np);
    T.py.resize(np);
    T.pz.resize(np);
    for (long k = 0; k < np; k++) {
      RD(R.reals(rv, 3));
      T.px[k] = rv[0];
      T.py[k] = rv[1];
      T.pz[k] = rv[2];
    }
  } else {
    return ifail(in, GFSPH_ERR_UNSUPPORTED,
                 path + ": ictop " + std::to_string(T.ictop) + " (0-7, -7, 10, 11, 15, 16 and +100 are read)");
  }
  // terrain properties (Topo_main 2:453-576): 1-4 the obsolete topo_props
  // table, read and unused by the propagation step; 10 the terrain zones
  RD(R.text());
  RD(R.ints(iv, 1));
  T.topo_props = (int)iv[0];
  const int itp = T.topo_props;
  if (itp == 11 || itp == 12)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": check dams (ic_topo_props " + std::to_string(itp) + ")");
  if (itp == 100 || itp == 101 || itp == -101)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": geological zones (ic_topo_props " + std::to_string(itp) + ")");
  if (itp >= 1 && itp <= 4) {
    const int nprops = itp == 1 ? 4 : (itp == 4 ? 7 : 6);
    RD(R.text());
    RD(R.ints(iv, 1));
    const long nz = iv[0];
    RD(R.text());
    for (long z = 0; z < nz; z++) RD(R.reals(rv, nprops));
  } else if (itp == 10) {
    RD(R.text());
    RD(R.ints(iv, 1));
    const long nz = iv[0];
    if (nz < 0 || nz > 32) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": more than 32 terrain zones");
    bool eros_pts = false;
    for (long z = 0; z < nz; z++) {
      gfsph_zone Z{};
      RD(R.text());
      RD(R.ints(iv, 1));
      Z.type = (int32_t)iv[0];
      RD(R.text());
      RD(R.ints(iv, 1));
      Z.law = (int32_t)iv[0];
      RD(R.text());
      RD(R.reals(rv, 6));
      for (int k = 0; k < 6; k++) Z.props[k] = rv[k];
      RD(R.text());
      RD(R.reals(rv, 10));
      for (int k = 0; k < 10; k++) Z.props[6 + k] = rv[k];
      const int g = (int)(Z.props[0] + 0.01);
      if (g == 102 || g == 200)
        return ifail(in, GFSPH_ERR_UNSUPPORTED,
                     path + ": terrain zone of geometry " + std::to_string(g) + " (structures, obstacles)");
      if (g == 101) eros_pts = true;
      T.zones.push_back(Z);
    }
    if (eros_pts) {  // Get_Eros_Pts reads <terrain>.eros.pts (2:943-978)
      std::string ep = path.substr(0, path.size() - 4) + ".eros.pts";
      FReader E;
      if (!E.open(ep)) return ifail(in, GFSPH_ERR_ARG, "cannot open " + ep);
      auto efail = [&]() { return ifail(in, GFSPH_ERR_ARG, ep + ": " + E.error()); };
      if (!E.text() || !E.reals(rv, 3)) return efail();
      const long npts = (long)rv[0];
      T.eros_type = (int)rv[1];
      T.eros_dxs = rv[2];
      if (npts < 0) return ifail(in, GFSPH_ERR_ARG, ep + ": npts_Eros");
      if (!E.text()) return efail();
      T.epts.resize(3 * npts);
      for (long k = 0; k < npts; k++) {
        if (!E.reals(rv, 3)) return efail();
        for (int j = 0; j < 3; j++) T.epts[3 * k + j] = rv[j];
      }
    }
  }
  return GFSPH_OK;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)); }
  hipError_t up(const T *h, size_t n) { return hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice); }
  hipError_t down(T *h, size_t n) const { return hipMemcpy(h, p, sizeof(T) * n, hipMemcpyDeviceToHost); }
};

#define HCHK(call)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return ifail(in, GFSPH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int blocks(long n) { return (int)((n + INP_BLOCK - 1) / INP_BLOCK); }

// CSR of point indices per cell, in point order (cells outside are skipped)
void bin_points(const std::vector<int> &cell, int ncell, std::vector<int> &start, std::vector<int> &pts) {
  start.assign(ncell + 1, 0);
  for (int c : cell)
    if (c >= 0) start[c + 1]++;
  for (int c = 0; c < ncell; c++) start[c + 1] += start[c];
  pts.assign(start[ncell], 0);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (size_t p = 0; p < cell.size(); p++)
    if (cell[p] >= 0) pts[fill[cell[p]]++] = (int)p;
}

// Get_DTM's grid (2:1227-1284) and the points' cells (2:1367-1373) for
// ictop 10, 11, 15, 16; ictop 0 takes the grid as read (2:224-252).  Host.
int grid_geometry(gfsph_inputs *in) {
  TopoRaw &T = in->topo;
  gfsph_grid &G = in->grid;
  G.fdiff = T.fdiff;
  if ((T.ictop >= 1 && T.ictop <= 7) || T.ictop == -7) {  // 2:262-290
    G.xmin = T.gx0;
    G.xmax = T.gx1;
    G.ymin = T.gy0;
    G.ymax = T.gy1;
    G.dx = (G.xmax - G.xmin) / (G.nx - 1);
    G.dy = (G.ymax - G.ymin) / (G.ny - 1);
    return GFSPH_OK;
  }
  if (T.ictop == 0) {
    double xmin = 1.e9, ymin = 1.e9, xmax = -1.e9, ymax = -1.e9;
    for (size_t k = 0; k < T.px.size(); k++) {
      if (T.px[k] <= xmin) xmin = T.px[k];
      if (T.px[k] >= xmax) xmax = T.px[k];
      if (T.py[k] <= ymin) ymin = T.py[k];
      if (T.py[k] >= ymax) ymax = T.py[k];
    }
    if (G.nx < 2 || G.ny < 2) return ifail(in, GFSPH_ERR_ARG, "terrain grid has fewer than 2 x 2 nodes");
    G.xmin = xmin;
    G.ymin = ymin;
    G.xmax = xmax;
    G.ymax = ymax;
    G.dx = (xmax - xmin) / (G.nx - 1);
    G.dy = (ymax - ymin) / (G.ny - 1);
    return GFSPH_OK;
  }
  const long np = (long)T.px.size();
  double xmin = 1.e9, ymin = 1.e9, xmax = -1.e9, ymax = -1.e9, zmax = -1.e9;
  for (long k = 0; k < np; k++) {
    if (T.px[k] <= xmin) xmin = T.px[k];
    if (T.px[k] >= xmax) xmax = T.px[k];
    if (T.py[k] <= ymin) ymin = T.py[k];
    if (T.py[k] >= ymax) ymax = T.py[k];
    if (T.pz[k] >= zmax) zmax = T.pz[k];
  }
  T.zmax = zmax;
  int nx = (int)((xmax - xmin) / T.deltx + 1), ny = (int)((ymax - ymin) / T.deltx + 1);
  double dxg = (xmax - xmin) / (nx - 1), dyg = (ymax - ymin) / (ny - 1);
  double xs = 0.0, ys = 0.0;
  if (T.ictop == 11) {
    xs = xmin;
    ys = ymin;
    for (long k = 0; k < np; k++) {
      T.px[k] = T.px[k] - xmin;
      T.py[k] = T.py[k] - ymin;
    }
    xmax = xmax - xmin;
    ymax = ymax - ymin;
    xmin = 0.0;
    ymin = 0.0;
  } else if (T.ictop == 15 || T.ictop == 16) {
    xs = xmin;
    ys = ymin;
    if (T.ictop == 15) {
      xmin = T.xw0 - xs;
      xmax = T.xw1 - xs;
      ymin = T.yw0 - ys;
      ymax = T.yw1 - ys;
    } else {
      xmin = T.xw0;
      xmax = T.xw1;
      ymin = T.yw0;
      ymax = T.yw1;
    }
    nx = (int)((xmax - xmin) / T.deltx + 1);
    ny = (int)((ymax - ymin) / T.deltx + 1);
    dxg = (xmax - xmin) / (nx - 1);
    dyg = (ymax - ymin) / (ny - 1);
    for (long k = 0; k < np; k++) {
      T.px[k] = T.px[k] - xs;
      T.py[k] = T.py[k] - ys;
    }
  }
  if (nx < 2 || ny < 2) return ifail(in, GFSPH_ERR_ARG, "terrain grid has fewer than 2 x 2 nodes");
  G.nx = nx;
  G.ny = ny;
  G.xmin = xmin;
  G.ymin = ymin;
  G.xmax = xmax;
  G.ymax = ymax;
  G.dx = dxg;
  G.dy = dyg;
  G.xshift = xs;
  G.yshift = ys;
  T.nbx = nx + 1;  // ipoigx runs 1 .. nx (nx + 1 if the division rounds up at the edge)
  T.nby = ny + 1;
  std::vector<int> cell(np, -1);
  for (long k = 0; k < np; k++) {
    const double xx = T.px[k], yy = T.py[k];
    const bool discr = (xx >= xmin) && (xx <= xmax) && (yy >= ymin) && (yy <= ymax);
    if (!discr) continue;
    const int ipx = (int)((xx - xmin) / dxg + 1), ipy = (int)((yy - ymin) / dyg + 1);
    const int bx = std::min(std::max(ipx - 1, 0), T.nbx - 1), by = std::min(std::max(ipy - 1, 0), T.nby - 1);
    cell[k] = by * T.nbx + bx;
  }
  bin_points(cell, T.nbx * T.nby, T.start, T.pts);
  return GFSPH_OK;
}

// Z of a generated grid (Topo_main 2:292-444, get_cone 2:1113-1151), node
// ipoig = iy nx + ix at (xmin + ix dx, ymin + iy dy); host, once.  The
// slope-with-transition (2) and bilinear (3) forms leave a node outside
// their segments at the previous node's value -- an input error here.
int generated_terrain(gfsph_inputs *in) {
  const TopoRaw &T = in->topo;
  const gfsph_grid &G = in->grid;
  const long np = (long)G.nx * G.ny;
  const std::vector<double> &g = T.gen;
  in->Z.assign(np, 0.0);
  in->in_model.assign(np, 1);
  for (long ip = 0; ip < np; ip++) {
    const double xx = G.xmin + (double)(ip % G.nx) * G.dx, yy = G.ymin + (double)(ip / G.nx) * G.dy;
    double zz = 0.0;
    if (T.ictop == 1) {
      zz = g[0];
    } else if (T.ictop == 2) {  // slope T1, circular transition of radius R, runout T3
      const double thetard = g[2] * 3.1416 / 180;
      const double x0 = 0.0, x1 = x0 + g[0], x2 = x1 + g[1] * sin(thetard), x3 = x2 + g[3];
      const double H1 = g[1] * (1. - cos(thetard));
      if (xx >= x0 && xx < x1) zz = H1 + (x1 - xx) * tan(thetard);
      else if (xx >= x1 && xx < x2) zz = g[1] * (1 - cos(asin((x2 - xx) / g[1])));
      else if (xx >= x2 && xx <= x3) zz = 0.0;
      else return ifail(in, GFSPH_ERR_ARG, "ictop 2: a node lies outside the slope's segments");
    } else if (T.ictop == 3) {  // bilinear in x through (x0, d0), (x1, d1), (x2, d2)
      const double x0 = g[0], d0 = g[1], x1 = g[2], d1 = g[3], x2 = g[4], d2 = g[5];
      if (xx >= x0 && xx <= x1) {
        const double xn0 = (xx - x1) / (x0 - x1), xn1 = 1. - xn0;
        zz = xn0 * d0 + xn1 * d1;
      } else if (xx >= x1 && xx <= x2) {
        const double xn1 = (xx - x2) / (x1 - x2), xn2 = 1. - xn1;
        zz = xn1 * d1 + xn2 * d2;
      } else {
        return ifail(in, GFSPH_ERR_ARG, "ictop 3: a node lies outside [x0, x2]");
      }
    } else if (T.ictop == 4) {  // multilinear in x; 0 outside the points
      const size_t nz = g.size() / 2;
      for (size_t k = 0; k + 1 < nz; k++) {
        const double x0 = g[k], x1 = g[k + 1];
        if (xx >= x0 && xx <= x1) {
          const double f0 = (xx - x1) / (x0 - x1), f1 = 1. - f0;
          zz = f0 * g[nz + k] + f1 * g[nz + k + 1];
          break;
        }
      }
    } else if (T.ictop == 5) {  // cone: r_top, R_bottom, Z_top, Z_bottom around the origin
      const double rr = sqrt(xx * xx + yy * yy);
      if (rr <= g[0]) zz = g[2];
      else if (rr > g[1]) zz = g[3];
      else zz = g[3] + (g[2] - g[3]) * (g[1] - rr) / (g[1] - g[0]);
    } else if (T.ictop == 6) {  // drum of radius R
      if (fabs(xx) > g[2]) return ifail(in, GFSPH_ERR_ARG, "ictop 6: a node lies outside the drum's radius");
      zz = g[2] * (1 - cos(asin(xx / g[2])));
    } else if (T.ictop == 7) {
      const double a0z = g[0] / (2. * g[1]);
      zz = a0z * xx * xx;
    } else {  // -7
      zz = 0.5 * g[0] * (1. + cos(xx * 3.141592 / g[1]));
    }
    in->Z[ip] = zz;
  }
  return GFSPH_OK;
}

// Get_DTM's gridding (2:1360-1404) on the device
int grid_build(gfsph_inputs *in) {
  TopoRaw &T = in->topo;
  const gfsph_grid &G = in->grid;
  const long npoig = (long)G.nx * G.ny;
  if (T.ictop == 0) {
    in->Z = T.pz;
    in->in_model.assign(npoig, 1);
    return GFSPH_OK;
  }
  if ((T.ictop >= 1 && T.ictop <= 7) || T.ictop == -7) return generated_terrain(in);
  const long np = (long)T.px.size();
  DevBuf<double> dx_, dy_, dz_, dZ;
  DevBuf<int> dstart, dpts, dmod;
  HCHK(dx_.alloc(np));
  HCHK(dy_.alloc(np));
  HCHK(dz_.alloc(np));
  HCHK(dstart.alloc(T.start.size()));
  HCHK(dpts.alloc(T.pts.size()));
  HCHK(dZ.alloc(npoig));
  HCHK(dmod.alloc(npoig));
  HCHK(dx_.up(T.px.data(), np));
  HCHK(dy_.up(T.py.data(), np));
  HCHK(dz_.up(T.pz.data(), np));
  HCHK(dstart.up(T.start.data(), T.start.size()));
  if (!T.pts.empty()) HCHK(dpts.up(T.pts.data(), T.pts.size()));
  const double distg = sqrt(G.dx * G.dx + G.dy * G.dy);
  Bins B{dstart.p, dpts.p, T.nbx, T.nby};
  k_dtm_idw<<<blocks(npoig), INP_BLOCK>>>(G.nx, G.ny, G.xmin, G.ymin, G.dx, G.dy, distg, dx_.p, dy_.p, dz_.p, B,
                                          T.zmax, dZ.p, dmod.p);
  HCHK(hipGetLastError());
  in->Z.resize(npoig);
  in->in_model.resize(npoig);
  HCHK(dZ.down(in->Z.data(), npoig));
  HCHK(dmod.down(in->in_model.data(), npoig));
  return GFSPH_OK;
}

// Get_Topol (2:1723-1939) with niter = 5 on the device
int topol_build(gfsph_inputs *in) {
  const gfsph_grid &G = in->grid;
  if (G.fdiff) return GFSPH_OK;  // finite differences on the device at create (Fdiff_derivs)
  const long npoig = (long)G.nx * G.ny;
  Mesh M{G.nx, G.ny, G.xmin, G.ymin, G.dx, G.dy};
  DevBuf<double> dZ, mml, T4, rhs, v0, v1, s, top9;
  HCHK(dZ.alloc(npoig));
  HCHK(mml.alloc(npoig));
  HCHK(T4.alloc(4 * npoig));
  HCHK(rhs.alloc(3 * npoig));
  HCHK(v0.alloc(3 * npoig));
  HCHK(v1.alloc(3 * npoig));
  HCHK(s.alloc(3 * npoig));
  HCHK(top9.alloc(9 * npoig));
  HCHK(dZ.up(in->Z.data(), npoig));
  const int nb = blocks(npoig);
  k_topo_mass<<<nb, INP_BLOCK>>>(M, mml.p);
  k_topo_pack<<<nb, INP_BLOCK>>>((int)npoig, dZ.p, nullptr, T4.p);
  // slopes (A2-A5)
  k_topo_rhs<<<nb, INP_BLOCK>>>(M, 0, T4.p, rhs.p);
  k_topo_first<<<nb, INP_BLOCK>>>((int)npoig, 2, rhs.p, mml.p, v0.p);
  double *a = v0.p, *b = v1.p;
  for (int it = 2; it <= 5; it++) {
    k_topo_iter<<<nb, INP_BLOCK>>>(M, 2, rhs.p, mml.p, a, b);
    std::swap(a, b);
  }
  HCHK(hipMemcpy(s.p, a, sizeof(double) * 3 * npoig, hipMemcpyDeviceToDevice));
  // curvatures (B1-B5)
  k_topo_pack<<<nb, INP_BLOCK>>>((int)npoig, dZ.p, s.p, T4.p);
  k_topo_rhs<<<nb, INP_BLOCK>>>(M, 1, T4.p, rhs.p);
  k_topo_first<<<nb, INP_BLOCK>>>((int)npoig, 3, rhs.p, mml.p, v0.p);
  a = v0.p;
  b = v1.p;
  for (int it = 2; it <= 5; it++) {
    k_topo_iter<<<nb, INP_BLOCK>>>(M, 3, rhs.p, mml.p, a, b);
    std::swap(a, b);
  }
  k_topo_final<<<nb, INP_BLOCK>>>((int)npoig, dZ.p, s.p, a, top9.p, nullptr);
  HCHK(hipGetLastError());
  in->topol.resize(9 * npoig);
  HCHK(top9.down(in->topol.data(), 9 * npoig));
  return GFSPH_OK;
}

// Read_Source_plus_New's reading, release grid and cells (6:5063-5168), np_Src > 0.  Host.
int read_pts7(gfsph_inputs *in, const std::string &path) {
  FReader R;
  if (!R.open(path)) return ifail(in, GFSPH_ERR_ARG, "cannot open " + path);
  SrcRaw &S = in->src;
  std::vector<double> rv;
  RD(R.text());
  RD(R.reals(rv, 5));
  const long nsrc = (long)rv[0];
  const double dx_s = rv[1];
  double d2x = rv[2], d2y = rv[3];
  S.facthsml = rv[4];
  S.delt2D = std::max(d2x, d2y);
  if (nsrc <= 0) return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": np_Src <= 0 (points from topol(15,:))");
  if (!(d2x > 0) || !(d2y > 0)) return ifail(in, GFSPH_ERR_ARG, path + ": grid spacing");
  RD(R.text());
  S.sx.resize(nsrc);
  S.sy.resize(nsrc);
  S.sh.resize(nsrc);
  double xmin = 1.e9, ymin = 1.e9, xmax = -1.e9, ymax = -1.e9;
  for (long k = 0; k < nsrc; k++) {
    RD(R.reals(rv, 3));
    S.sx[k] = rv[0] - in->grid.xshift;
    S.sy[k] = rv[1] - in->grid.yshift;
    S.sh[k] = rv[2];
    if (S.sx[k] <= xmin) xmin = S.sx[k];
    if (S.sx[k] >= xmax) xmax = S.sx[k];
    if (S.sy[k] <= ymin) ymin = S.sy[k];
    if (S.sy[k] >= ymax) ymax = S.sy[k];
  }
  S.nx = (int)((xmax - xmin) / d2x + 1);
  S.ny = (int)((ymax - ymin) / d2y + 1);
  if (S.nx < 2 || S.ny < 2) return ifail(in, GFSPH_ERR_ARG, path + ": release grid has fewer than 2 x 2 nodes");
  d2x = (xmax - xmin) / (S.nx - 1);
  d2y = (ymax - ymin) / (S.ny - 1);
  S.xmin = xmin;
  S.ymin = ymin;
  S.d2x = d2x;
  S.d2y = d2y;
  S.dist2D = sqrt(d2x * d2x + d2y * d2y);
  S.ndx = (int)std::max(2., dx_s / d2x);
  S.ndy = (int)std::max(2., dx_s / d2y);
  if ((2 * S.ndx + 1) * (2 * S.ndy + 1) > MAX_WINDOW)
    return ifail(in, GFSPH_ERR_UNSUPPORTED, path + ": point spacing too large for the release grid");
  S.nbx = S.nx + 1;
  S.nby = S.ny + 1;
  std::vector<int> cell(nsrc);
  for (long k = 0; k < nsrc; k++) {
    const int ipx = (int)((S.sx[k] - xmin) / d2x + 1), ipy = (int)((S.sy[k] - ymin) / d2y + 1);
    cell[k] = std::min(std::max(ipy - 1, 0), S.nby - 1) * S.nbx + std::min(std::max(ipx - 1, 0), S.nbx - 1);
  }
  bin_points(cell, S.nbx * S.nby, S.start, S.pts);
  return GFSPH_OK;
}

// Read_Source_plus_New's gridding (6:5170-5247) on the device; nodes with
// weight become particles in node order
int release_build(gfsph_inputs *in) {
  const SrcRaw &S = in->src;
  const long nsrc = (long)S.sx.size(), np2 = (long)S.nx * S.ny;
  DevBuf<double> dx_, dy_, dh, aux;
  DevBuf<int> dstart, dpts;
  HCHK(dx_.alloc(nsrc));
  HCHK(dy_.alloc(nsrc));
  HCHK(dh.alloc(nsrc));
  HCHK(dstart.alloc(S.start.size()));
  HCHK(dpts.alloc(S.pts.size()));
  HCHK(aux.alloc(4 * np2));
  HCHK(dx_.up(S.sx.data(), nsrc));
  HCHK(dy_.up(S.sy.data(), nsrc));
  HCHK(dh.up(S.sh.data(), nsrc));
  HCHK(dstart.up(S.start.data(), S.start.size()));
  HCHK(dpts.up(S.pts.data(), S.pts.size()));
  Bins B{dstart.p, dpts.p, S.nbx, S.nby};
  k_src_idw<<<blocks(np2), INP_BLOCK>>>(S.nx, S.ny, S.xmin, S.ymin, S.d2x, S.d2y, S.dist2D, S.ndx, S.ndy, dx_.p,
                                        dy_.p, dh.p, B, aux.p);
  k_src_area<<<blocks(np2), INP_BLOCK>>>(S.nx, S.ny, aux.p);
  HCHK(hipGetLastError());
  std::vector<double> ah(4 * np2);
  HCHK(aux.down(ah.data(), 4 * np2));
  const double dArea = S.d2x * S.d2y / 4;
  in->n = 0;
  for (long k = 0; k < np2; k++) {
    if (!(ah[4 * k + 1] >= 1.e-6)) continue;
    const int ix = (int)(k % S.nx), iy = (int)(k / S.nx);
    double rho = ah[4 * k + 2] / ah[4 * k + 1];
    if (rho < in->P.h_inf) rho = in->P.h_inf;
    in->x.push_back(S.xmin + (double)ix * S.d2x);
    in->x.push_back(S.ymin + (double)iy * S.d2y);
    in->v.push_back(0.0);
    in->v.push_back(0.0);
    in->h.push_back(rho);
    in->mass.push_back(dArea * rho * ah[4 * k + 3]);
    in->hsml.push_back(S.delt2D * S.facthsml);
    in->u.push_back(0.0);
    in->itype.push_back(1);
    in->n++;
  }
  return GFSPH_OK;
}

// Get_Restart_Info_SW (6:3268-3340).  Host.
int read_restart(gfsph_inputs *in, const std::string &path) {
  FReader R;
  if (!R.open(path)) return ifail(in, GFSPH_ERR_ARG, "cannot open " + path);
  std::vector<double> rv;
  RD(R.text());
  RD(R.reals(rv, 2));
  const long n = (long)rv[0];
  in->run.time_restart = rv[1];
  if (n <= 0) return ifail(in, GFSPH_ERR_ARG, path + ": npoin_s");
  RD(R.text());
  for (long k = 0; k < n; k++) {
    RD(R.reals(rv, 10));
    in->itype.push_back((int32_t)rv[1]);
    in->x.push_back(rv[2]);
    in->x.push_back(rv[3]);
    in->v.push_back(rv[4]);
    in->v.push_back(rv[5]);
    in->h.push_back(rv[6]);
    in->mass.push_back(rv[7]);
    in->hsml.push_back(rv[8]);
    in->u.push_back(rv[9]);
  }
  in->n = (int)n;
  return GFSPH_OK;
}

// everything that needs no device: the files and the grids' geometry
int parse_all(gfsph_inputs *in, const std::string &dir, const std::string &name, const std::string &topo) {
  const std::string base = dir.empty() ? std::string() : dir + "/";
  int rc = read_master(in, base + name + ".MASTER.dat");
  if (rc) return rc;
  std::string src;
  if ((rc = read_dat(in, base + name + ".dat", src))) return rc;
  if (in->run.present) {
    if (in->run.t_integ_alg != 4)
      return ifail(in, GFSPH_ERR_UNSUPPORTED, "SPH_t_Integ_Alg " + std::to_string(in->run.t_integ_alg) +
                                                  " (4, RK4, is the implemented scheme)");
    in->P.dt = in->run.dt_sph;
  }
  if ((rc = read_top(in, base + topo + ".top", in->topo))) return rc;
  if ((rc = grid_geometry(in))) return rc;
  // the release file is <problem name>.pts (6:5053-5055); a restart is restart.dat
  if (in->icunk == 7) return read_pts7(in, base + name + ".pts");
  if (in->icunk == 1 || in->icunk == 11 || in->icunk == 12 || in->icunk == 13) return GFSPH_OK;  // generated
  return read_restart(in, base + "restart.dat");
}

// Terrain zones after the grid (Topo_main reads them after Get_DTM and
// before the slopes, 2:453-617): each geometry-101 zone runs Get_Eros_Pts,
// which sets topol(15,:) afresh and may change Z (type_Eros -1, -2, -3).
int zones_build(gfsph_inputs *in) {
  TopoRaw &T = in->topo;
  if (T.zones.empty()) return GFSPH_OK;
  const gfsph_grid &G = in->grid;
  const long np = (long)G.nx * G.ny;
  in->eros_avail.assign(np, 0.0);
  bool any101 = false;
  std::vector<int32_t> listed(np);
  for (size_t z = 0; z < T.zones.size(); z++) {
    gfsph_zone &Z = T.zones[z];
    if ((int)(Z.props[0] + 0.01) != 101) continue;
    if (!any101) in->zone_nodes.assign(np, 0);
    any101 = true;
    const bool expl = T.ictop == 0;  // explicit grid: coorg are the coordinates read
    double box[4];
    const int rc = gfsph_terrain_eros_pts(G.nx, G.ny, G.xmin, G.ymin, G.dx, G.dy, expl ? T.px.data() : nullptr,
                                          expl ? T.py.data() : nullptr, (int32_t)(T.epts.size() / 3),
                                          T.epts.data(), T.eros_type, T.eros_dxs, in->Z.data(),
                                          in->eros_avail.data(), listed.data(), box);
    if (rc) return ifail(in, rc, "terrain zone " + std::to_string(z + 1) + ": erodible points");
    for (int k = 0; k < 4; k++) Z.props[1 + k] = box[k];  // basal_props(2:5) (2:989-992)
    for (long k = 0; k < np; k++)
      if (listed[k]) in->zone_nodes[k] |= (int32_t)(1u << z);
  }
  return GFSPH_OK;
}

int build_all(gfsph_inputs *in) {
  int rc;
  if ((rc = grid_build(in)) || (rc = zones_build(in)) || (rc = topol_build(in))) return rc;
  if (in->icunk == 7 && (rc = release_build(in))) return rc;
  in->built = true;
  return GFSPH_OK;
}

}  // namespace

extern "C" {

int gfsph_inputs_parse(const char *dir, const char *name, const char *topo, gfsph_inputs **out) {
  if (!out || !name) return GFSPH_ERR_ARG;
  *out = nullptr;
  gfsph_inputs *in = new (std::nothrow) gfsph_inputs();
  if (!in) return GFSPH_ERR_ARG;
  *out = in;
  return parse_all(in, dir ? dir : "", name, topo ? topo : name);
}

int gfsph_inputs_read(const char *dir, const char *name, const char *topo, gfsph_inputs **out) {
  int rc = gfsph_inputs_parse(dir, name, topo, out);
  if (rc) return rc;
  return build_all(*out);
}

int gfsph_inputs_info(gfsph_inputs *in, gfsph_params *p, gfsph_run *run, gfsph_grid *grid, int32_t *npart) {
  if (!in) return GFSPH_ERR_ARG;
  if (p) *p = in->P;
  if (run) *run = in->run;
  if (grid) *grid = in->grid;
  if (npart) *npart = in->n;
  return GFSPH_OK;
}

int gfsph_inputs_dem(gfsph_inputs *in, double *Z, int32_t *in_model, double *topol) {
  if (!in) return GFSPH_ERR_ARG;
  if (!in->built) return ifail(in, GFSPH_ERR_ARG, "the terrain is built by gfsph_inputs_read");
  const size_t np = in->Z.size();
  if (Z) memcpy(Z, in->Z.data(), sizeof(double) * np);
  if (in_model) memcpy(in_model, in->in_model.data(), sizeof(int32_t) * np);
  if (topol) {
    if (in->topol.empty()) return ifail(in, GFSPH_ERR_ARG, "slopes are finite differences (ictop >= 100)");
    memcpy(topol, in->topol.data(), sizeof(double) * 9 * np);
  }
  return GFSPH_OK;
}

int gfsph_inputs_particles(gfsph_inputs *in, double *x, double *v, double *h, double *mass, double *hsml,
                           double *u, int32_t *itype) {
  if (!in) return GFSPH_ERR_ARG;
  if (in->n <= 0) return ifail(in, GFSPH_ERR_ARG, "no release particles (a gridded release is built by gfsph_inputs_read)");
  const size_t n = in->n;
  if (x) memcpy(x, in->x.data(), sizeof(double) * 2 * n);
  if (v) memcpy(v, in->v.data(), sizeof(double) * 2 * n);
  if (h) memcpy(h, in->h.data(), sizeof(double) * n);
  if (mass) memcpy(mass, in->mass.data(), sizeof(double) * n);
  if (hsml) memcpy(hsml, in->hsml.data(), sizeof(double) * n);
  if (u) memcpy(u, in->u.data(), sizeof(double) * n);
  if (itype) memcpy(itype, in->itype.data(), sizeof(int32_t) * n);
  return GFSPH_OK;
}

int gfsph_inputs_create(gfsph_inputs *in, gfsph_ctx **out) {
  if (!in || !out) return GFSPH_ERR_ARG;
  if (!in->built || in->n <= 0) return ifail(in, GFSPH_ERR_ARG, "inputs not built (gfsph_inputs_read)");
  *out = nullptr;
  if (!(in->P.dt > 0)) return ifail(in, GFSPH_ERR_ARG, "dt_sph is not set (no MASTER.dat): set it in the params");
  gfsph_dem d{};
  d.nx = in->grid.nx;
  d.ny = in->grid.ny;
  d.xmin = in->grid.xmin;
  d.ymin = in->grid.ymin;
  d.dx = in->grid.dx;
  d.dy = in->grid.dy;
  d.xmax = in->grid.xmax;
  d.ymax = in->grid.ymax;
  d.Z = in->Z.data();
  d.in_model = in->in_model.data();
  d.topol = in->topol.empty() ? nullptr : in->topol.data();
  d.nzones = (int32_t)in->topo.zones.size();
  d.zones = in->topo.zones.empty() ? nullptr : in->topo.zones.data();
  d.zone_nodes = in->zone_nodes.empty() ? nullptr : in->zone_nodes.data();
  d.eros_avail = in->eros_avail.empty() ? nullptr : in->eros_avail.data();
  int rc = gfsph_create(&in->P, &d, in->n, in->x.data(), in->v.data(), in->h.data(), in->mass.data(),
                        in->hsml.data(), in->u.data(), in->itype.data(), out);
  if (rc) {
    in->err = std::string("gfsph_create: ") + (*out ? gfsph_last_error(*out) : "failed");
    if (*out) gfsph_destroy(*out);
    *out = nullptr;
    return rc;
  }
  if (in->run.present && in->run.adapt_dt) {
    if ((rc = gfsph_set_adaptive_dt(*out, in->run.adapt_dt))) {
      in->err = gfsph_last_error(*out);
      gfsph_destroy(*out);
      *out = nullptr;
      return rc;
    }
  }
  if (in->run.time_restart != 0.0) gfsph_set_time(*out, in->run.time_restart);
  return GFSPH_OK;
}

int gfsph_inputs_cases(gfsph_inputs *in, int32_t *mode, int32_t *ncases, int32_t *nproblems, double *table,
                       double *stop3) {
  if (!in) return GFSPH_ERR_ARG;
  const CaseTable &T = in->cases;
  if (mode) *mode = T.mode;
  if (ncases) *ncases = T.ncases;
  if (nproblems) *nproblems = (int32_t)T.OC.size();
  if (table)
    for (size_t p = 0; p < T.OC.size(); p++)
      memcpy(table + p * T.OC[p].size(), T.OC[p].data(), sizeof(double) * T.OC[p].size());
  if (stop3) {
    stop3[0] = T.vstop;
    stop3[1] = T.timestop;
    stop3[2] = T.dtimestop;
  }
  return GFSPH_OK;
}

// Every problem of the case table as one realisation of a batch: the release
// and const(:) changed as Input_SW changes them for the problem
// (6:1040-1096 particles: 501 depth and mass factor, 511 expansion about the
// centroid, 521 shift; 6:1208-1232 const(:), index 2 read as a porosity when
// <= 1.01), ic_adapt_dt_sph, and the case control of the loop the reference
// runs for each problem (Driver_Flow for mode 1: no plot-time cut;
// Driver_Flow_23 for mode 2: cut at every plot_step*dt).
int gfsph_inputs_batch_create(gfsph_inputs *in, int32_t pixels, gfsph_ctx **out) {
  if (!in || !out) return GFSPH_ERR_ARG;
  *out = nullptr;
  const CaseTable &T = in->cases;
  if (!in->built || in->n <= 0) return ifail(in, GFSPH_ERR_ARG, "inputs not built (gfsph_inputs_read)");
  if (T.mode == 0 || T.OC.empty()) return ifail(in, GFSPH_ERR_ARG, "MASTER.dat holds no case table");
  if (!(in->P.dt > 0)) return ifail(in, GFSPH_ERR_ARG, "dt_sph is not set");
  const int R = (int)T.OC.size(), n0 = in->n;
  const long ntot = (long)R * n0;
  if (ntot > 0x7fffffffL) return ifail(in, GFSPH_ERR_ARG, "too many particles over all problems");
  std::vector<double> cst((size_t)R * 111), x(2 * ntot), v(2 * ntot), h(ntot), mass(ntot), hsml(ntot), u(ntot);
  std::vector<int32_t> counts(R, n0);
  for (int p = 0; p < R; p++) {
    double *c = &cst[(size_t)p * 111];
    memcpy(c, in->P.cst, sizeof(double) * 111);
    double *px = &x[2L * p * n0], *pv = &v[2L * p * n0], *ph = &h[(long)p * n0], *pm = &mass[(long)p * n0];
    memcpy(px, in->x.data(), sizeof(double) * 2 * n0);
    memcpy(pv, in->v.data(), sizeof(double) * 2 * n0);
    memcpy(ph, in->h.data(), sizeof(double) * n0);
    memcpy(pm, in->mass.data(), sizeof(double) * n0);
    memcpy(&hsml[(long)p * n0], in->hsml.data(), sizeof(double) * n0);
    memcpy(&u[(long)p * n0], in->u.data(), sizeof(double) * n0);
    for (int ic = 0; ic < T.ncases; ic++) {
      if (T.mode == 1 && ic != (int)T.OC[p][0] - 1) continue;  // mode 1: the varied variable only
      const int idx = (int)(T.C[ic][0] + 0.01);
      const double cv = T.OC[p][3 + ic];
      if (idx == 501) {
        for (int i = 0; i < n0; i++) {
          ph[i] = ph[i] * cv;
          pm[i] = pm[i] * cv;
        }
      } else if (idx == 511) {
        double cg[2] = {0.0, 0.0};
        for (int i = 0; i < n0; i++) {
          cg[0] = cg[0] + px[2 * i];
          cg[1] = cg[1] + px[2 * i + 1];
          if (T.mode == 1) pm[i] = pm[i] * cv;  // the mode-1 branch scales the masses in both loops (6:1094-1101)
        }
        cg[0] = cg[0] / n0;
        cg[1] = cg[1] / n0;
        for (int i = 0; i < n0; i++) {
          px[2 * i] = cg[0] + cv * (px[2 * i] - cg[0]);
          px[2 * i + 1] = cg[1] + cv * (px[2 * i + 1] - cg[1]);
          pm[i] = pm[i] * cv;
        }
      } else if (idx == 521) {
        const double dx = cv, dy = T.OC[p][3 + T.ncases + 1];
        for (int i = 0; i < n0; i++) {
          px[2 * i] = px[2 * i] + dx;
          px[2 * i + 1] = px[2 * i + 1] + dy;
        }
      } else if (idx == 401) {
        // the initial pore pressure's factor, applied to the created context
        // below (gfsph_batch_scale_pwp)
      } else if (idx >= 1 && idx <= 110) {
        if (T.mode >= 2 && idx == 2 && cv <= 1.01) {  // porosity given: mixture density (6:1216-1226)
          if (!(c[17] >= 1.e-4))
            return ifail(in, GFSPH_ERR_ARG, "a porosity for const(2) needs the solid and water densities const(17, 18)");
          c[2] = (double)(float)((1. - cv) * c[17] + cv * c[18]);
        } else {
          c[idx] = (double)(float)cv;
        }
      } else {
        return ifail(in, GFSPH_ERR_UNSUPPORTED, "case variable index " + std::to_string(idx));
      }
    }
  }
  gfsph_dem d{};
  d.nx = in->grid.nx;
  d.ny = in->grid.ny;
  d.xmin = in->grid.xmin;
  d.ymin = in->grid.ymin;
  d.dx = in->grid.dx;
  d.dy = in->grid.dy;
  d.xmax = in->grid.xmax;
  d.ymax = in->grid.ymax;
  d.Z = in->Z.data();
  d.in_model = in->in_model.data();
  d.topol = in->topol.empty() ? nullptr : in->topol.data();
  d.nzones = (int32_t)in->topo.zones.size();
  d.zones = in->topo.zones.empty() ? nullptr : in->topo.zones.data();
  d.zone_nodes = in->zone_nodes.empty() ? nullptr : in->zone_nodes.data();
  d.eros_avail = in->eros_avail.empty() ? nullptr : in->eros_avail.data();
  int rc = gfsph_batch_create(&in->P, &d, R, counts.data(), cst.data(), x.data(), v.data(), h.data(), mass.data(),
                              hsml.data(), u.data(), out);
  auto bail = [&](const char *what) {
    in->err = std::string(what) + ": " + (*out ? gfsph_last_error(*out) : "failed");
    if (*out) gfsph_destroy(*out);
    *out = nullptr;
    return rc;
  };
  if (rc) return bail("gfsph_batch_create");
  // variable 401 (6:1453-1479, pore pressure only): u and Uaux of problem p
  // times its value, one variable after another; in mode 1 only where it is
  // the problem's varied variable
  for (int ic = 0; ic < T.ncases && in->P.icpwp != 0; ic++) {
    if ((int)(T.C[ic][0] + 0.01) != 401) continue;
    std::vector<double> fac(R, 1.0);
    for (int p = 0; p < R; p++)
      if (!(T.mode == 1 && ic != (int)T.OC[p][0] - 1)) fac[p] = T.OC[p][3 + ic];
    if ((rc = gfsph_batch_scale_pwp(*out, fac.data()))) return bail("gfsph_batch_scale_pwp");
  }
  if (in->run.present && in->run.adapt_dt && (rc = gfsph_set_adaptive_dt(*out, in->run.adapt_dt)))
    return bail("gfsph_set_adaptive_dt");
  gfsph_case_ctl ctl;
  ctl.vstop = T.vstop;
  ctl.timestop = T.timestop;
  ctl.dtimestop = T.dtimestop;
  ctl.dt = in->run.dt;
  ctl.time_end = in->run.time_end > 0 ? in->run.time_end : 1.e30;
  ctl.time_plot = T.mode >= 2 ? in->run.plot_step * in->run.dt : 0.0;
  ctl.pixels = pixels && T.mode >= 2 ? 1 : 0;
  ctl.pad_ = 0;
  if ((rc = gfsph_batch_cases(*out, &ctl))) return bail("gfsph_batch_cases");
  return GFSPH_OK;
}

int gfsph_terrain_eros_pts(int32_t nx, int32_t ny, double xmin, double ymin, double dx, double dy, const double *cx,
                           const double *cy, int32_t npts, const double *pts, int32_t type_eros, double dx_s,
                           double *Z, double *avail, int32_t *listed, double *box) {
  if (nx < 2 || ny < 2 || npts < 0 || (npts > 0 && !pts) || !Z || !avail || !listed || !(dx > 0) || !(dy > 0))
    return GFSPH_ERR_ARG;
  const long np = (long)nx * ny;
  auto coorg = [&](long ip, int d) {  // node ip (0-based)
    if (d == 0) return cx ? cx[ip] : xmin + (double)(ip % nx) * dx;
    return cy ? cy[ip] : ymin + (double)(ip / nx) * dy;
  };
  const double dist2D = sqrt(dx * dx + dy * dy);
  std::vector<double> a2(np, 0.0), a3(np, 0.0);
  for (long k = 0; k < np; k++) avail[k] = 0.0;  // topol(15,:) = 0
  double xemin = 1.e10, xemax = -1.e10, yemin = 1.e10, yemax = -1.e10;
  for (int32_t e = 0; e < npts; e++) {
    const double xe = pts[3 * e], ye = pts[3 * e + 1];
    if (xe >= xemax) xemax = xe;
    if (ye >= yemax) yemax = ye;
    if (xe <= xemin) xemin = xe;
    if (ye <= yemin) yemin = ye;
  }
  if (box) {
    box[0] = xemin;
    box[1] = xemax;
    box[2] = yemin;
    box[3] = yemax;
  }
  const int ixmin = (int)((xemin - xmin) / dx + 1), ixmax = (int)((xemax - xmin) / dx + 1);
  const int iymin = (int)((yemin - ymin) / dy + 1), iymax = (int)((yemax - ymin) / dy + 1);
  const int ndx = (int)std::max(2., dx_s / dx), ndy = (int)std::max(2., dx_s / dy);
  for (int32_t e = 0; e < npts; e++) {
    const double xx = pts[3 * e], yy = pts[3 * e + 1], dep = pts[3 * e + 2];
    const int ipx = (int)((xx - xmin) / dx + 1), ipy = (int)((yy - ymin) / dy + 1);
    const int inx0 = std::max(std::max(1, ipx - ndx), ixmin), inx1 = std::min(std::min(nx, ipx + ndx + 1), ixmax);
    const int iny0 = std::max(std::max(1, ipy - ndy), iymin), iny1 = std::min(std::min(ny, ipy + ndy + 1), iymax);
    for (int inx = inx0; inx <= inx1; inx++)
      for (int iny = iny0; iny <= iny1; iny++) {
        const long ip = (long)(iny - 1) * nx + (inx - 1);
        double dx2 = xx - coorg(ip, 0);
        dx2 = dx2 * dx2;
        double dy2 = yy - coorg(ip, 1);
        dy2 = dy2 * dy2;
        double dist = std::max(1.e-6, sqrt(dx2 + dy2));
        dist = dist / dist2D;
        a2[ip] = a2[ip] + 1.0 / dist;
        a3[ip] = a3[ip] + dep / dist;
      }
  }
  for (long ip = 0; ip < np; ip++) {
    listed[ip] = a2[ip] >= 1.e-6;
    if (listed[ip]) avail[ip] = a3[ip] / a2[ip];
  }
  if (type_eros == -1) {
    for (long ip = 0; ip < np; ip++) Z[ip] = Z[ip] + avail[ip];
  } else if (type_eros == -2) {
    for (long ip = 0; ip < np; ip++) {
      const double z0 = Z[ip];
      Z[ip] = std::max(Z[ip], avail[ip]);
      avail[ip] = Z[ip] - z0;
    }
  } else if (type_eros == -3) {
    for (long ip = 0; ip < np; ip++) Z[ip] = Z[ip] - avail[ip];
  }
  return GFSPH_OK;
}

int gfsph_inputs_zones(gfsph_inputs *in, int32_t *nzones, gfsph_zone *zones, int32_t *zone_nodes,
                       double *eros_avail) {
  if (!in) return GFSPH_ERR_ARG;
  if (!in->built) return ifail(in, GFSPH_ERR_ARG, "inputs not built (gfsph_inputs_read)");
  const long np = (long)in->grid.nx * in->grid.ny;
  if (nzones) *nzones = (int32_t)in->topo.zones.size();
  if (zones && !in->topo.zones.empty())
    memcpy(zones, in->topo.zones.data(), sizeof(gfsph_zone) * in->topo.zones.size());
  if (zone_nodes) {
    for (long k = 0; k < np; k++) zone_nodes[k] = in->zone_nodes.empty() ? 0 : in->zone_nodes[k];
  }
  if (eros_avail) {
    for (long k = 0; k < np; k++) eros_avail[k] = in->eros_avail.empty() ? 0.0 : in->eros_avail[k];
  }
  return GFSPH_OK;
}

int gfsph_inputs_resample(gfsph_inputs *in, uint64_t seed) {
  if (!in) return GFSPH_ERR_ARG;
  CaseTable &T = in->cases;
  if (!(T.mode == 3 || T.mode == 4 || T.mode == 5 || T.mode == 31 || T.mode == 32))
    return ifail(in, GFSPH_ERR_ARG, "MASTER.dat holds no sampled case set (ic_cases_mat 3, 4, 31, 32)");
  T.seed = seed;
  mc_fill(T);
  return GFSPH_OK;
}

const char *gfsph_inputs_error(gfsph_inputs *in) { return in ? in->err.c_str() : "null inputs"; }

void gfsph_inputs_destroy(gfsph_inputs *in) { delete in; }

}  // extern "C"
