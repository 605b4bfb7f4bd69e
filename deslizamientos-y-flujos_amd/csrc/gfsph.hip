// gfsph.hip -- MI355X (gfx950) implementation of the Geoflow-SPH depth
// integrated SPH propagation step behind the C ABI of include/gfsph.h.
//
// One step (frozen pair geometry, "7 SPH_MAIN_2019.f90":784-799):
//   k_grid         grid of the active bounding box (cell = half the largest
//                  support radius)
//   k_keys         cell key per particle + atomic count per cell
//   k_scan_*       exclusive scan of the cell counts -> cell starts (and
//                  the counts cleared for the next build)
//   k_scatter      counting-sort scatter
//   k_reorder      gather the particle state into cell order, each cell
//                  ordered by previous slot (deterministic order)
//   k_density      summed depth h_i = W(0)m_i + sum m_j w_ij  (5a:1763)
//   k_pforce       pressure-force gather (IntForces_SW, 6:6669)
//   k_rk           every RK stage per particle: DEM bilinear lookup
//                  (2:2395), gravity, basal rheology and limiter
//                  (6:9355-9901), erosion (6:8664-8812), stage update
//                  (7:808-922), pore pressure (6:11176-11310), domain flag
//                  (7:1764), h update (7:2866), next-step bounding box
//                  (Bingham: per-stage kernels)
//   k_erode_nodes  terrain update of erosion (6:8790-8806)
//   k_xsph         XSPH velocity correction (7:3042-3099, 7:987-993)
// All arithmetic is FP64.  The file is compiled with -ffp-contract=off so
// expressions round as the reference's FP64 build writes them.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <functional>
#include <mutex>
#include <type_traits>
#include <vector>

#include "gfsph.h"

#define GF_BLOCK 256

// ---------------------------------------------------------------------------
// Device-side parameter blocks
// ---------------------------------------------------------------------------

struct DevParams {
  int nfric, if_frictional, if_V2byR, icpwp, nuaux, pa_sph, skf, scale_k, sle, av_vel, npoin0;
  int nstages;
  double dt, h_inf, cgra, cgra05, dens, xi_voellmy, constMu_CF, K_Bagnold;
  double tauy0, constK, visco, tanfi8, hfrict0, Bfact, comp, alphaS, hrelpw2, c73;
  double fact_dens, fact_pwp_max, fact_pwp_min, fpwp_max, fpwp_min;
  double c3, c6, c7, c8, c47;
  double f1[4], f2[4], fin;
  double pi;
  double kc;  // kernel normalisation: 15/(7 pi) for skf 1, 7/(4 pi) for skf 4
  // erosion (ic_erosion): law const(21) (1 Hungr, 2 slope angle, 3 basal
  // stress), its constants const(22..25), grain and fluid densities
  // const(17), const(18), the bottom-velocity factor from const(47);
  // eros_single: the realisation holds one particle, whose depth then grows
  // with the eroded depth too (6:8756-8758)
  int eros_on, eros_law, eros_single, pad_e;
  // icpwp = 2: column of npwp nodes (0: no column), its bed / surface
  // conditions and the consolidation coefficient Cv = 4 Bfact / pi^2
  int npwp, bc_t1, bc_t2, pad_c;
  double bc_v1, bc_v2, Cv;
  // pressure re-evaluated in every stage: g along the terrain normal
  // (ic_cosTheta = 1) and / or K0 active / passive (ic_K0 = 1)
  int costheta, ic_k0;
  double k0a, k0p;
  double keros, eros_c23, eros_c24, eros_c25, denss, densw_e, dens_bar_eff, slip_fact, avail0;
  // nfric = 23 (exact Herschel-Bulkley root): hb_m = 1 (m = 1), 2 (m = 2) or
  // 3 (m = 1/3 or 1/2, a_hat = 4.5 K |hv| / (tau_y h^2)); hb_A = (1/Cm)
  // (K/tau_y)^(1/m), the run-constant factor of a_hat for m = 1 and 2
  int hb_m, pad_h;
  double hb_A;
  // terrain zones of geometry 1 read the realisation's current Z at
  // DevDem::zcur + zoff (topol(1), which erosion lowers)
  long zoff;
};

// One terrain zone (basal_type, basal_law, basal_props of 2:516-576): geom =
// int(props(1) + 0.1) as Which_Topo_Zone takes it, egeom = int(props(1) +
// 0.01) as Get_Erosion_Rate_SW does; cth, sth: cos and sin of the geometry-5
// angle, from the host's libm.
struct DevZone {
  int type, law, geom, egeom;
  double p[16];
  double cth, sth;
};

struct DevDem {
  int nx, ny;
  double xmin, ymin, dx, dy, xmax, ymax;
  const double2 *zxy;  // [npoig] (Zx, Zy)
  const double *top;   // [npoig][6] top4..top9, or null
  // terrain zones: zkinds bit 0 = a friction zone exists, bit 1 = an
  // erosion zone, bit 2 = a pore-pressure zone; znodes: node-list bits of geometry-101 zones; zcur, inm:
  // Z and poig_in_model for geometry 1 (null when no zone needs them)
  int nzones, zkinds;
  const DevZone *zones;
  const int *znodes;
  const double *zcur;
  const int *inm;
  const double *zavail;  // topol(15): available erosion depth (ErosGrid::avail), per realisation at zoff
};

struct GridInfo {
  double xmin, ymin, cell;
  double hmax;    // largest smoothing length of the active particles
  int ncx, ncy, ncell;
  int uniform_h;  // every active particle has the same smoothing length
};

// Cells are half the largest support radius, so a particle's candidates lie
// in the 5 x 5 cells around its own (rows cy-2 .. cy+2).
#ifndef NROWS
#define NROWS 5
#endif
#define RAD ((NROWS - 1) / 2)
#define CELL_FRAC (1.0 / RAD)
// cells per particle the grids may use (the same area per particle for any RAD)
#define CAP_PER_PART (4 * RAD * RAD)

// bounding-box accumulator of active particles (ordered-integer encoding)
struct BBoxAcc {
  unsigned long long xmin, xmax, ymin, ymax, hmax, hmin;
  unsigned int nactive, pad;
};

// particle state in cell order; nb packs what neighbour gathers read
// (x, y, hsml, mass) into one 32-byte record
struct SortedSet {
  double4 *nb;
  double2 *pos, *vel;
  double *h, *mass, *hsml, *u, *uaux;
  int *flag, *id;
  int *rid;     // realisation of the particle (0 for a single run)
  double *col;  // icpwp = 2: pore-pressure columns [npwp][n] (node-major), else null
  const double *mass0;  // nfric = 3: masses before any erosion by caller id (mass0, 5:41), else null
};

// Cell index of one or several independent realisations: realisation r has
// its own grid (bounding box of its particles) and owns the cell keys
// [coff[r], coff[r+1]); out-of-domain particles take key coff[R].
struct Cells {
  const GridInfo *grid;  // [R]
  const int *coff;       // [R+1]
  const int *ntot;       // coff + R: number of cells of all realisations
  const int *start;      // [ntot+2] first slot of every key
};

// step outputs in the same (cell) order as the sorted set that produced them
struct NewSet {
  double2 *pos, *vel;
  double *h, *hsml, *u, *uaux;
  int *flag;
  double *col;  // icpwp = 2: columns after the step [npwp][n]
  double *dh;   // icpwp = 2: the step's depth change Drho_RK, read by the column update
};

// The state after a step: positions, depths, smoothing lengths, pore
// pressures and flags are in the step's output set; velocities (after XSPH),
// masses and caller ids are in the sorted set, both in the same slot order.
struct CurView {
  const double2 *pos, *vel;
  const double *h, *mass, *hsml, *u, *uaux;
  const int *flag, *id, *rid;
};

__device__ __forceinline__ unsigned long long ord_of(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double val_of(unsigned long long o) {
  unsigned long long b = (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
  double d;
  memcpy(&d, &b, sizeof(d));
  return d;
}

// ---------------------------------------------------------------------------
// a2. kernel_TOOLS ("5a My_SPH_Tools_2019.f90":2062-2180), ndimn = 2
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sph_kernel(int skf, double pi, double r, double dx, double dy,
                                           double hsml, double &w, double &gx, double &gy) {
  double q = r / hsml, factor, a;
  w = 0.0;
  gx = 0.0;
  gy = 0.0;
  if (skf == 1) {
    factor = 15.e0 / (7.e0 * pi * hsml * hsml);
    if (q >= 0 && q <= 1.e0) {
      w = factor * (2. / 3. - q * q + q * q * q / 2.);
      a = factor * (-2. + 3. / 2. * q) / (hsml * hsml);
      gx = a * dx;
      gy = a * dy;
    } else if (q > 1.e0 && q <= 2) {
      double t = 2. - q;
      w = factor * 1.e0 / 6.e0 * (t * t * t);
      a = -(factor * 1.e0 / 6.e0 * 3. * (t * t) / hsml);
      gx = a * (dx / r);
      gy = a * (dy / r);
    }
  } else if (skf == 2) {
    factor = 1.e0 / (hsml * hsml * pi);
    if (q >= 0 && q <= 3) {
      w = factor * exp(-q * q);
      gx = w * (-2. * dx / hsml / hsml);
      gy = w * (-2. * dy / hsml / hsml);
    }
  } else if (skf == 3) {
    factor = 7.e0 / (478.e0 * pi * hsml * hsml);
    double t3 = 3 - q, t2 = 2 - q, t1 = 1 - q;
    double t3_4 = t3 * t3 * t3 * t3, t2_4 = t2 * t2 * t2 * t2;
    if (q >= 0 && q <= 1) {
      w = factor * (t3_4 * t3 - 6 * t2_4 * t2 + 15 * (t1 * t1 * t1 * t1 * t1));
      a = factor * ((-120 + 120 * q - 50 * q * q) / (hsml * hsml));
      gx = a * dx;
      gy = a * dy;
    } else if (q > 1 && q <= 2) {
      w = factor * (t3_4 * t3 - 6 * t2_4 * t2);
      a = factor * (-5 * t3_4 + 30 * t2_4) / hsml;
      gx = a * (dx / r);
      gy = a * (dy / r);
    } else if (q > 2 && q <= 3) {
      w = factor * (t3_4 * t3);
      a = factor * (-5 * t3_4) / hsml;
      gx = a * (dx / r);
      gy = a * (dy / r);
    }
  } else {  // skf == 4, Wendland
    factor = 7.e0 / (4.e0 * pi * hsml * hsml);
    if (q >= 0 && q <= 2) {
      double b = 1 - q / 2;
      w = factor * ((b * b * b * b) * (2 * q + 1));
      a = factor * (-5 * (b * b * b)) * ((1 / hsml / hsml));
      gx = a * dx;
      gy = a * dy;
    }
  }
}

// Reciprocal for kernel evaluation: the hardware estimate refined by two
// Newton-Raphson steps (relative error a few 2^-53 for normal x).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// r = sqrt(dr) and ir = 1/sqrt(dr) for kernel evaluation: the hardware
// reciprocal-root estimate refined by two Newton-Raphson steps and one
// correction of the root; dr = 0 gives r = ir = 0.
__device__ __forceinline__ void root_nr(double dr, double &r, double &ir) {
  double y = __builtin_amdgcn_rsq(dr);
  const double hx = 0.5 * dr;
  double e = fma(-(hx * y), y, 0.5);
  y = fma(y, e, y);
  e = fma(-(hx * y), y, 0.5);
  y = fma(y, e, y);
  double g = dr * y;
  const double d = fma(-g, g, dr);
  g = fma(d, 0.5 * y, g);
  const bool pos = dr > 0.0;
  r = pos ? g : 0.0;
  ir = pos ? y : 0.0;
}

// Kernel value and gradient for a compile-time kernel type from the squared
// separation dr, arranged around one reciprocal of the smoothing length (and
// one reciprocal root where the gradient needs dx/r) instead of the
// reference's six divisions; kc is the normalisation 15/(7 pi) (cubic
// spline) or 7/(4 pi) (Wendland) computed once.  Values agree with
// kernel_TOOLS to a few ulps; only the neighbour predicate must round
// exactly as the reference, and it is evaluated separately.
// UH: ihu is 1/hsml already (all smoothing lengths equal, so hsml is the
// particle's own and its reciprocal is computed once per particle).
template <int SKF, bool UH = false>
__device__ __forceinline__ void sph_kernel_t(double kc, double pi, double dr, double dx, double dy, double hsml,
                                             double &w, double &gx, double &gy, double ihu = 0.0) {
  if (SKF == 1) {
    double r, ir;
    root_nr(dr, r, ir);
    const double ih = UH ? ihu : rcp_nr(hsml);
    const double ih2 = ih * ih;
    const double q = r * ih;
    const double factor = kc * ih2;
    const double w1 = factor * (2. / 3. - q * q + 0.5 * (q * q * q));
    const double a1 = factor * (-2. + 1.5 * q) * ih2;
    const double t = 2. - q;
    const double w2 = factor * (1.0 / 6.0) * (t * t * t);
    const double a2 = -(factor * 0.5 * (t * t) * ih) * ir;
    const bool in1 = q <= 1.0, in2 = q <= 2.0;
    w = in1 ? w1 : (in2 ? w2 : 0.0);
    const double a = in1 ? a1 : (in2 ? a2 : 0.0);
    gx = a * dx;
    gy = a * dy;
  } else if (SKF == 4) {
    double r, ir;
    root_nr(dr, r, ir);
    const double ih = UH ? ihu : rcp_nr(hsml);
    const double ih2 = ih * ih;
    const double q = r * ih;
    const double factor = kc * ih2;
    const double b = 1 - 0.5 * q;
    const double b3 = b * b * b;
    const bool in = q <= 2.0;
    w = in ? factor * ((b3 * b) * (2 * q + 1)) : 0.0;
    const double a = in ? factor * (-5 * b3) * ih2 : 0.0;
    gx = a * dx;
    gy = a * dy;
  } else {
    sph_kernel(SKF, pi, sqrt(dr), dx, dy, hsml, w, gx, gy);
  }
}

// kernel value at r = 0 (self contribution in sum_density_Tools, 5a:1818-1822)
__device__ __forceinline__ double sph_self_w(int skf, double pi, double hsml) {
  double w, a, b;
  sph_kernel(skf, pi, 0.0, 0.0, 0.0, hsml, w, a, b);
  return w;
}

// a1. pair predicate of grid_find_NEW_tools3 (5a:1010-1019), rounded exactly
// as the reference writes it: dr = dx*dx + dy*dy without contraction.
// pair_pred_fast first rejects, without the square root, every candidate
// with dr > t*t*(1+2^-46), t = scale_k*mh: then sqrt(dr) > t even after the
// roundings of t*t and of the square root, so the exact test would fail too;
// all other candidates take the exact test.
__device__ __forceinline__ bool pair_pred(int scale_k, double2 pi_, double2 pj, double hi, double hj,
                                          double &dx, double &dy, double &r, double &mh) {
  dx = __dsub_rn(pi_.x, pj.x);
  dy = __dsub_rn(pi_.y, pj.y);
  double dr = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
  r = sqrt(dr);
  mh = __dadd_rn(hi, hj) / 2.;
  return r < scale_k * mh;
}

__device__ __forceinline__ bool pair_pred_fast(int scale_k, double2 pi_, double2 pj, double hi, double hj) {
  double dx = __dsub_rn(pi_.x, pj.x);
  double dy = __dsub_rn(pi_.y, pj.y);
  double dr = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
  double mh = __dadd_rn(hi, hj) / 2.;
  double t = scale_k * mh;
  if (dr > __dmul_rn(__dmul_rn(t, t), 1.0 + 0x1p-46)) return false;
  return sqrt(dr) < t;
}

// ---------------------------------------------------------------------------
// a7. Get_Z_Topo ("2 SPH_FEM_topo_2019.f90":2395-2491)
// ---------------------------------------------------------------------------
struct TopoPt {
  double Zx, Zy, t4, t5, t6, t7, t8, t9;
};

__device__ __forceinline__ void get_z_topo(const DevDem &d, double xx, double yy, bool want_top,
                                           TopoPt &o) {
  int ipx = (int)((xx - d.xmin) / d.dx + 1);
  int ipy = (int)((yy - d.ymin) / d.dy + 1);
  if (ipx == d.nx && fabs(d.xmax - xx) <= 1.e-3) ipx = ipx - 1;
  if (ipy == d.ny && fabs(d.ymax - yy) <= 1.e-3) ipy = ipy - 1;
  // in-bounds guard: the reference only warns here (2:2421-2435)
  ipx = ipx < 1 ? 1 : (ipx > d.nx - 1 ? d.nx - 1 : ipx);
  ipy = ipy < 1 ? 1 : (ipy > d.ny - 1 ? d.ny - 1 : ipy);
  long n1 = (long)(ipy - 1) * d.nx + (ipx - 1);
  long n2 = n1 + 1, n4 = n1 + d.nx, n3 = n4 + 1;
  double xref = d.xmin + (ipx - 1) * d.dx;
  double yref = d.ymin + (ipy - 1) * d.dy;
  double xi = (xx - xref) / d.dx, eta = (yy - yref) / d.dy;
  double sh1 = (1. - xi) * (1. - eta), sh2 = xi * (1. - eta), sh3 = xi * eta, sh4 = (1. - xi) * eta;
  double2 a = d.zxy[n1], b = d.zxy[n2], c = d.zxy[n3], e = d.zxy[n4];
  o.Zx = sh1 * a.x + sh2 * b.x + sh3 * c.x + sh4 * e.x;
  o.Zy = sh1 * a.y + sh2 * b.y + sh3 * c.y + sh4 * e.y;
  if (want_top) {
    const double *T1 = d.top + 6 * n1, *T2 = d.top + 6 * n2, *T3 = d.top + 6 * n3, *T4 = d.top + 6 * n4;
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = sh1 * T1[k] + sh2 * T2[k] + sh3 * T3[k] + sh4 * T4[k];
    o.t4 = v[0];
    o.t5 = v[1];
    o.t6 = v[2];
    o.t7 = v[3];
    o.t8 = v[4];
    o.t9 = v[5];
  }
}

// Z at (xx, yy) as Get_Z_Topo interpolates topol(1) (2:2395-2470), with
// the weighting by poig_in_model at the edge of the model (2:2455-2462);
// zc: the realisation's current Z.
__device__ __forceinline__ double z_topo(const DevDem &d, const double *zc, double xx, double yy) {
  int ipx = (int)((xx - d.xmin) / d.dx + 1);
  int ipy = (int)((yy - d.ymin) / d.dy + 1);
  if (ipx == d.nx && fabs(d.xmax - xx) <= 1.e-3) ipx = ipx - 1;
  if (ipy == d.ny && fabs(d.ymax - yy) <= 1.e-3) ipy = ipy - 1;
  ipx = ipx < 1 ? 1 : (ipx > d.nx - 1 ? d.nx - 1 : ipx);
  ipy = ipy < 1 ? 1 : (ipy > d.ny - 1 ? d.ny - 1 : ipy);
  const long n1 = (long)(ipy - 1) * d.nx + (ipx - 1);
  const long n2 = n1 + 1, n4 = n1 + d.nx, n3 = n4 + 1;
  const double xref = d.xmin + (ipx - 1) * d.dx;
  const double yref = d.ymin + (ipy - 1) * d.dy;
  const double xi = (xx - xref) / d.dx, eta = (yy - yref) / d.dy;
  const double sh1 = (1. - xi) * (1. - eta), sh2 = xi * (1. - eta), sh3 = xi * eta, sh4 = (1. - xi) * eta;
  double Zp = sh1 * zc[n1] + sh2 * zc[n2] + sh3 * zc[n3] + sh4 * zc[n4];
  if (d.inm) {
    const int i1 = d.inm[n1], i2 = d.inm[n2], i3 = d.inm[n3], i4 = d.inm[n4];
    if (i1 + i2 + i3 + i4 < 4) {
      Zp = i1 * sh1 * zc[n1] + i2 * sh2 * zc[n2] + i3 * sh3 * zc[n3] + i4 * sh4 * zc[n4];
      const double den = i1 * sh1 + i2 * sh2 + i3 * sh3 + i4 * sh4;
      if (den > 1.e-3) Zp = Zp / den;
    }
  }
  return Zp;
}

// Which_Topo_Zone (2:2635-2808): the first zone of basal type btype that
// holds (xx, yy), or -1.  Geometry 101 takes the cell of the point (the
// corner node's index as Get_Z_Topo forms it, nodes n1, n1+1, n1+1+nx,
// n1+nx) and is inside when one of the four nodes is in the zone's list.
__device__ __noinline__ int which_zone(const DevDem &d, const DevParams &P, double xx, double yy, int btype) {
  for (int z = 0; z < d.nzones; z++) {
    const DevZone &Z = d.zones[z];
    if (Z.type != btype) continue;
    bool in;
    if (Z.geom <= 3) {
      const double Zp = Z.geom == 1 ? z_topo(d, d.zcur + P.zoff, xx, yy) : (Z.geom == 2 ? xx : yy);
      in = (Z.p[2] - Zp) * (Zp - Z.p[1]) >= 0;
    } else if (Z.geom == 4) {
      in = (Z.p[2] - xx) * (xx - Z.p[1]) >= 0 && (Z.p[4] - yy) * (yy - Z.p[3]) >= 0;
    } else if (Z.geom == 5) {
      const double Rvx = xx - Z.p[1], Rvy = yy - Z.p[2];
      const double Ra = Z.cth * Rvx + Z.sth * Rvy;
      const double Rb = (-Z.sth) * Rvx + Z.cth * Rvy;
      in = Ra >= 0.0 && Ra <= Z.p[3] && Rb >= 0.0 && Rb <= Z.p[4];
    } else {  // 101
      int ipx = (int)((xx - d.xmin) / d.dx + 1);
      int ipy = (int)((yy - d.ymin) / d.dy + 1);
      if (ipx == d.nx && fabs(d.xmax - xx) <= 1.e-3) ipx = ipx - 1;
      if (ipy == d.ny && fabs(d.ymax - yy) <= 1.e-3) ipy = ipy - 1;
      const long np = (long)d.nx * d.ny;
      const long n1 = (long)(ipy - 1) * d.nx + (ipx - 1);
      const long nn[4] = {n1, n1 + 1, n1 + 1 + d.nx, n1 + d.nx};
      in = false;
      for (int k = 0; k < 4; k++)
        if (nn[k] >= 0 && nn[k] < np && ((d.znodes[nn[k]] >> z) & 1)) in = true;
    }
    if (in) return z;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// a6 + a8. ExtForces_SW(0) body for one particle (6:8532-8657) with
// Get_Grav_Ext_SW (6:9241-9349) and basal_friction_Ext_SW (6:9355-9901).
// ---------------------------------------------------------------------------
// Bingham intermediate regime (6:9609-9618): returns true and Tau0 when
// 0.01 <= a_hat <= 50.
__device__ __forceinline__ bool bingham_tau0(const DevParams &P, double hvmod, double hfrict, double &Tau0) {
  double a_hat = (6. * P.constK * hvmod) / (P.tauy0 * (hfrict * hfrict));
  if (a_hat > 50 || a_hat < 0.01) return false;
  double a2 = 48., b2 = -32. * a_hat - 114., c2 = 65.;
  double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
  Tau0 = P.tauy0 / xh;
  return true;
}

// candidate Tau0 of one particle for the last-valid scan (NaN if the
// particle is not in the intermediate regime); same arithmetic as ext_force
__device__ __forceinline__ double bingham_candidate(const DevParams &P, double rho_i, double vx1, double vy1) {
  double h = rho_i, hvmod;
  if (h < P.comp) {
    h = P.comp;
    hvmod = 0;
  } else {
    double hu = h * vx1, hv = h * vy1;
    hvmod = sqrt(hu * hu + hv * hv);
  }
  double Tau0;
  if (bingham_tau0(P, hvmod, fmax(h, P.hfrict0), Tau0)) return Tau0;
  return __longlong_as_double(0x7ff8000000000000ll);
}

// Get_Erosion_Rate_SW (6:10289-10404), main-region laws, single phase.  h,
// vmod, vxx, vyy, tanfi and frictx/fricty are the values ExtForces_SW's
// shared variables hold after basal_friction_Ext_SW.
__device__ __forceinline__ double erosion_rate(const DevParams &P, double h, double vmod, double vxx, double vyy,
                                               double Zpx, double Zpy, double tanfi, double frictx,
                                               double fricty) {
  double er = 0.0;
  if (P.eros_law == 1) {
    er = P.keros * h * vmod;
  } else if (P.eros_law == 2) {
    const double denss = P.denss, densw = P.densw_e;
    const double vol_conc = (P.dens - densw) / (denss - densw);
    if (vmod > 1.e-5) {
      const double cosa = vxx / vmod, sina = vyy / vmod;
      const double gradZv = Zpx * cosa + Zpy * sina;
      const double theta = atan2(-gradZv, 1.);
      double xnum = (denss - densw) * vol_conc;
      const double xden = xnum + densw;
      xnum = xnum * tanfi;
      const double thetae = atan(xnum / xden);
      if (theta > thetae) er = P.keros * P.eros_c23 * vmod * tan(theta - thetae);
    }
  } else if (P.eros_law == 3) {
    if (vmod > 1.e-5) {
      const double Taub = P.dens * sqrt(frictx * frictx + fricty * fricty);  // dens_bar = dens (6:9007)
      const double Tauf = P.eros_c23 + P.dens_bar_eff * P.cgra * h * (1. - P.eros_c25) * P.eros_c24;
      if (Taub > Tauf && vmod > 1.e-3) er = P.keros * (Taub - Tauf) / P.dens_bar_eff;
    }
  }
  return er;
}

// Erosion zones of the terrain (6:10407-10503): inside an erosion zone its
// own law replaces er -- note the zone laws' own forms: law 2 takes the
// equilibrium angle as atan2(xden, xnum), law 3 the basal stress dens h
// |frict| against c + dens g h (1 - beta) tan(phi_b); a zone of geometry 101
// or 102 erodes only where its cell's nodes have depth available
// (topol(15), the realisation's current values).
__device__ __noinline__ double zone_erosion_rate(const DevParams &P, const DevDem &dem, double xx, double yy,
                                                 double er, double h, double vmod, double vxx, double vyy,
                                                 double Zpx, double Zpy, double tanfi, double frictx, double fricty) {
  const int z = which_zone(dem, P, xx, yy, 2);
  if (z < 0) return er;
  const DevZone &Z = dem.zones[z];
  if (Z.law == 0) {
    er = 0.0;
  } else if (Z.law == 1) {
    er = Z.p[6] * h * vmod;
  } else if (Z.law == 2) {
    const double denss = P.denss, densw = P.densw_e;
    const double vol_conc = (P.dens - densw) / (denss - densw);
    er = 0.0;
    if (vmod > 1.e-5) {  // (at rest the reference stops: 6:10445-10447)
      const double cosa = vxx / vmod, sina = vyy / vmod;
      const double gradZv = Zpx * cosa + Zpy * sina;
      const double theta = atan2(-gradZv, 1.);
      double xnum = (denss - densw) * vol_conc;
      const double xden = xnum + densw;
      xnum = xnum * tanfi;
      const double thetae = atan2(xden, xnum);
      if (theta > thetae) er = Z.p[6] * Z.p[7] * vmod * tan(theta - thetae);
    }
  } else if (Z.law == 3) {
    er = 0.0;
    if (vmod > 1.e-5) {
      const double Taub = P.dens * h * sqrt(frictx * frictx + fricty * fricty);
      const double Tauf = Z.p[7] + P.dens * P.cgra * h * (1. - Z.p[9]) * Z.p[8];
      if (Taub > Tauf && vmod > 1.e-3) er = Z.p[6] * (Taub - Tauf) / P.dens;
    }
  }
  if (Z.egeom == 101 || Z.egeom == 102) {
    const int ipx = (int)((xx - dem.xmin) / dem.dx + 1);
    const int ipy = (int)((yy - dem.ymin) / dem.dy + 1);
    const long np = (long)dem.nx * dem.ny;
    const long n1 = (long)(ipy - 1) * dem.nx + (ipx - 1);
    const long nn[4] = {n1, n1 + 1, n1 + dem.nx, n1 + dem.nx + 1};
    const double *av = dem.zavail + P.zoff;
    double auxt = 0.0;
    for (int k = 0; k < 4; k++) auxt = auxt + ((nn[k] >= 0 && nn[k] < np) ? av[nn[k]] : 0.0);
    if (auxt < 1.e-6) er = 0.0;
  }
  return er;
}

// tau_in: for nfric = 4, the Tau0 resolved by the last-valid scan (the
// particle's own value when it is in the intermediate regime).
// EROS: erosion is compiled in.  Inside the RK stages (end = false) the
// erosion rate adds its momentum term (6:8664-8693); after them (end = true,
// ExtForces_SW(1)) exdvxdt is not cleared, so (acx, acy) -- the particle's
// exdvxdt of the last stage -- is where the gravity term is added and what
// the limiter reads, and only the rate is wanted (er).
// LAW = 5: compiled for nfric = 5 alone (frictional with optional Voellmy
// turbulence, no curvature term, g vertical) -- every realisation of the
// context has that law, so the other laws' code and registers are left out;
// LAW = 0: every law, chosen at run time.
#define LAW_ANY 0
#define LAW_FRIC5 5
// nsoil: (mass - mass0)/mass of the particle, read by nfric = 3 alone.
template <bool EROS = false, int LAW = LAW_ANY>
__device__ __forceinline__ void ext_force(const DevParams &P, const DevDem &dem, double xx, double yy,
                                          double vx1, double vy1, double rho_i, double &u_i,
                                          double uaux1, double inx, double iny, double tau_in, double nsoil,
                                          double &exo,
                                          double &eyo, double &duo, double acx = 0.0, double acy = 0.0,
                                          bool end = false, double *er_out = nullptr) {
  const double dt = P.dt;
  double h, hu, hv, hinv, vmod, hvmod, hvmodinv;
  double vxx = vx1, vyy = vy1;
  double gravz = P.cgra;  // g along the terrain normal with ic_cosTheta = 1 (6:9264-9270)
  double frictx = 0, fricty = 0, TotFrict = 0, cf5 = 0, hfrict, pwp;
  double xi_voellmy = P.xi_voellmy;
  double exx = (EROS && end) ? acx : 0.0, exy = (EROS && end) ? acy : 0.0, du = 0.0;
  double tanfi_m = 0.0;  // ExtForces_SW's tanfi after basal_friction_Ext_SW

  h = rho_i;
  hu = h * vx1;
  hv = h * vy1;
  if (h < P.comp) {
    h = P.comp;
    hu = 0;
    hv = 0;
    hinv = 0;
    vmod = 0;
    hvmod = 0;
    hvmodinv = 0;
  } else {
    hinv = 1 / h;
    hvmod = sqrt(hu * hu + hv * hv);
    vmod = hvmod * hinv;
    hvmodinv = (hvmod >= 1.e-5) ? 1. / hvmod : 0.0;
  }

  TopoPt tp;
  // the curvature terms are interpolated only when the law uses them (V^2/R);
  // the reference interpolates them for every frictional law (6:9257-9259)
  constexpr bool F5 = LAW == LAW_FRIC5;
  get_z_topo(dem, xx, yy, !F5 && P.if_V2byR != 0, tp);
  if (!F5 && P.costheta) {
    const double deng = sqrt(1. + tp.Zx * tp.Zx + tp.Zy * tp.Zy);
    gravz = P.cgra / deng;
    exx = exx + (-P.cgra * tp.Zx / deng);
    exy = exy + (-P.cgra * tp.Zy / deng);
  } else {
    exx = exx + (-P.cgra * tp.Zx);
    exy = exy + (-P.cgra * tp.Zy);
  }

  if (P.icpwp == 1) pwp = u_i;
  else if (P.icpwp == 2 || P.icpwp == 3) pwp = uaux1;  // Uaux(1,:), the bed value
  else pwp = 0.0;
  hfrict = fmax(h, P.hfrict0);
  const int nfric = F5 ? 5 : P.nfric;

  if (!F5 && nfric == 1) {
    double cf1 = P.cgra * xi_voellmy * xi_voellmy;
    double hp = pow(hfrict, P.c73);
    frictx = cf1 * hu * hvmod / hp;
    fricty = cf1 * hv * hvmod / hp;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && nfric == 2) {
    double cf2a = P.constK / (8. * P.dens);
    double h2 = (hfrict * hfrict);
    frictx = cf2a * hu / h2;
    fricty = cf2a * hv / h2;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && nfric == 3) {
    // Bingham erosion (6:9443-9488): Bingham on the eroded fraction
    // nsoil = (m - m0)/m, bounded below by Manning's law with n = const(8)
    const double fsoilB = 1. - exp(-P.c3 * nsoil);
    const double constKB = P.constK * fsoilB, tauy0B = P.tauy0 * fsoilB;
    const double densB = 1000. + (P.dens - 1000.) * nsoil;
    if (nsoil > 0.01) {
      const double h2 = hfrict * hfrict;
      const double a_hat = (6. * constKB * hvmod) / (tauy0B * h2);
      if (a_hat > 50) {
        const double cf4b = 3 * constKB / densB;
        frictx = cf4b * hu / h2;
        fricty = cf4b * hv / h2;
      } else if (a_hat < 0.01) {
        const double cf4b = tauy0B / densB;
        frictx = cf4b * hu * hvmodinv;
        fricty = cf4b * hv * hvmodinv;
      } else {
        const double a2 = 48., b2 = -32. * a_hat - 114., c2 = 65.;
        const double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
        const double cf4b = (tauy0B / xh) / densB;
        frictx = cf4b * hu * hvmodinv;
        fricty = cf4b * hv * hvmodinv;
      }
    }
    const double cf1 = P.cgra * P.c8 * P.c8;
    const double hp = pow(hfrict, P.c73);
    const double fxx = cf1 * hu * hvmod / hp, fyy = cf1 * hv * hvmod / hp;
    if (fabs(fxx) > fabs(frictx)) frictx = fxx;
    if (fabs(fyy) > fabs(fricty)) fricty = fyy;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && nfric == 23) {
    // Herschel-Bulkley with the exact root of the regime polynomial
    // (6:9538-9545, 6:9576-9605); Tau0 is the particle's own (not carried)
    const double h2 = hfrict * hfrict;
    double a_hat, a2 = 48., b2, c2 = 65.;
    if (P.hb_m == 3) a_hat = (4.5 * P.constK * hvmod) / (P.tauy0 * h2);
    else a_hat = P.hb_A * (hvmod / h2);
    if (P.hb_m == 2) {
      a2 = 10.;
      b2 = -a_hat - 19.2188;
      c2 = 9.1875;
    } else {
      b2 = -32. * a_hat - 114.;
    }
    const double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
    const double cf4b = (P.tauy0 / xh) / P.dens;
    frictx = cf4b * hu * hvmodinv;
    fricty = cf4b * hv * hvmodinv;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && (nfric == 4 || nfric == 24 || nfric == 25)) {
    const double Tauy0 = P.tauy0, constK = P.constK, m_vexp = P.visco;
    if (nfric == 4) {
      // every regime ends with cf4b = Tau0/dens (6:9619-9622); Tau0 is this
      // particle's own in the intermediate regime, else the carried one
      double Tau0 = tau_in;
      bingham_tau0(P, hvmod, hfrict, Tau0);
      double cf4b = Tau0 / P.dens;
      frictx = cf4b * hu * hvmodinv;
      fricty = cf4b * hv * hvmodinv;
      (void)constK;
    } else {
      double Tau0 = P.constK * pow(vmod / hfrict, m_vexp);
      double Cm = (nfric == 24) ? 1. : pow((2 * m_vexp + 1) / m_vexp, m_vexp);
      Tau0 = Tauy0 + Tau0 * Cm;
      double cf4b = Tau0 / P.dens;
      frictx = cf4b * hu * hvmodinv;
      fricty = cf4b * hv * hvmodinv;
    }
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (F5 || P.if_frictional == 1) {
    double cv2byR = 0.0, tanfi;
    if (!F5 && P.if_V2byR == 1) {
      vxx = hu * hinv;
      vyy = hv * hinv;
      vmod = sqrt(vxx * vxx + vyy * vyy);
      if (vmod >= 1e-5) {
        double cosa = vxx / vmod, sina = vyy / vmod;
        double cos2 = cosa * cosa, sin2 = sina * sina, cbs = sina * cosa;
        double xnum = tp.t7 * cos2 + tp.t8 * cbs + tp.t9 * sin2;
        double xden = tp.t4 * cos2 + tp.t5 * cbs + tp.t6 * sin2;
        cv2byR = vmod * vmod * (xnum / xden);
      }
    }
    tanfi = P.tanfi8;
    if (!F5 && (nfric == 8 || nfric == 26 || nfric == 27)) {
      xi_voellmy = 0.0;
      double Froude = vmod / sqrt(gravz * h);
      if (nfric == 8 || nfric == 26) {
        double ff = (1 + P.c3 * pow(Froude * P.c6 / h, P.c7));
        tanfi = tanfi * ff;
      } else {
        double Inert = Froude * P.c6 / h;
        tanfi = tanfi + (P.c3 - tanfi) / (1.0 + (P.c7 / Inert));
      }
    }
    double hhh = rho_i;
    if (pwp > P.fpwp_max * hhh) pwp = P.fpwp_max * hhh;
    if (pwp < P.fpwp_min * hhh) pwp = P.fpwp_min * hhh;
    u_i = pwp;  // 6:9675
    if (P.icpwp == 1 || P.icpwp == 2 || P.icpwp == 3) {
      double factpw = 1. - pwp * hinv;
      if (factpw < P.fact_pwp_min) factpw = P.fact_pwp_min;
      if (factpw > P.fact_pwp_max) factpw = P.fact_pwp_max;
      tanfi = tanfi * factpw;
    }
    if (!F5 && (dem.zkinds & 1)) {  // friction zones of the terrain (6:9697-9702, 10515-10545)
      const int z = which_zone(dem, P, xx, yy, 1);
      if (z >= 0 && dem.zones[z].p[6] >= 0) tanfi = dem.zones[z].p[6];
    }
    tanfi_m = tanfi;
    double gtanfi = (gravz + cv2byR) * tanfi * P.fact_dens;
    if (gtanfi < 0.0) gtanfi = 0.0;
    cf5 = gtanfi;
    frictx = cf5 * hu * hvmodinv * h;
    fricty = cf5 * hv * hvmodinv * h;
    if (xi_voellmy > 0.0) {
      vxx = vx1;
      vyy = vy1;
      vmod = vxx * vxx + vyy * vyy;
      vmod = (vmod > 0.0) ? sqrt(vmod) : 0.0;
      frictx = frictx + P.cgra * vmod * vxx / xi_voellmy;
      fricty = fricty + P.cgra * vmod * vyy / xi_voellmy;
    } else if (xi_voellmy < 0.) {
      vxx = vx1;
      vyy = vy1;
      vmod = vxx * vxx + vyy * vyy;
      vmod = (vmod > 0.0) ? sqrt(vmod) : 0.0;
      frictx = frictx + P.constMu_CF * vmod * vxx / (hfrict * hfrict);
      fricty = fricty + P.constMu_CF * vmod * vyy / (hfrict * hfrict);
    }
    if (!F5 && (nfric == 9 || nfric == 10)) {
      double C67 = P.c6 + P.c7;
      double frict_abs = sqrt(frictx * frictx + fricty * fricty);
      double fact9 = pow(2. * P.c8 * vmod / hfrict, 1. / C67) * pow(frict_abs, -P.c6 / C67);
      if (P.c47 > 1.e-6) fact9 = fact9 * ((1 - P.c47) / (1 + P.c47));
      frictx = frictx * (1 + fact9);
      fricty = fricty * (1 + fact9);
    }
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && nfric == 11) {
    vxx = hu * hinv;
    vyy = hv * hinv;
    frictx = P.K_Bagnold * vxx * vmod / (hfrict * hfrict);
    fricty = P.K_Bagnold * vyy * vmod / (hfrict * hfrict);
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (!F5 && nfric == 22) {
    // simplified Bingham + power law (6:9803-9816), as written there: a
    // negative velocity component adds nothing and the y term goes to frictx
    double tb_abs = P.tauy0 + (P.constK * pow(vmod / hfrict, P.visco));
    tb_abs = tb_abs / P.dens;
    if (vxx >= 1.e-4) frictx = tb_abs * vxx / vmod;
    if (vyy >= 1.e-4) frictx = frictx + tb_abs * vyy / vmod;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  }

  // limiter (6:9836-9895)
  double dcx = -frictx * hinv, dcy = -fricty * hinv;
  double ax = exx + inx, ay = exy + iny;
  if (vmod > 1.e-5) {
    double vsx = vx1 + dt * (ax + dcx), vsy = vy1 + dt * (ay + dcy);
    double vdot = vsx * vx1 + vsy * vy1;
    if (vdot < 0) {
      double tx = vx1 / vmod, ty = vy1 / vmod;
      double accelT = ax * tx + ay * ty;
      double deccelT = dcx * tx + dcy * ty;
      if (accelT * deccelT < 0) {
        dcx = -vx1 / dt - ax;
        dcy = -vy1 / dt - ay;
        frictx = -dcx * h;
        fricty = -dcy * h;
      }
    }
  } else {
    double accelT = sqrt(ax * ax + ay * ay);
    if (accelT > 1.e-5) {
      double ctf = TotFrict;
      if (F5 || P.if_frictional == 1) ctf = cf5;
      frictx = ctf * ax / accelT;
      fricty = ctf * ay / accelT;
      double fricT = sqrt(frictx * frictx + fricty * fricty);
      if (accelT <= fricT) {
        frictx = ax;
        fricty = ay;
      }
    } else {
      frictx = 0.0;
      fricty = 0.0;
    }
    frictx = frictx * h;
    fricty = fricty * h;
  }
  if (P.icpwp == 1) {
    double betaw = P.Bfact * hinv * hinv / P.hrelpw2;
    du = -(u_i / dt) * (1 - exp(-betaw * dt));
  }
  exo = exx - frictx / h;
  eyo = exy - fricty / h;
  if (EROS && P.eros_on) {
    double er = erosion_rate(P, h, vmod, vxx, vyy, tp.Zx, tp.Zy, tanfi_m, frictx, fricty);
    if (dem.zkinds & 2) er = zone_erosion_rate(P, dem, xx, yy, er, h, vmod, vxx, vyy, tp.Zx, tp.Zy, tanfi_m, frictx, fricty);
    if (!end) {
      exo = exo - er * vx1 * hinv * P.slip_fact;
      eyo = eyo - er * vy1 * hinv * P.slip_fact;
    }
    if (er_out) *er_out = er;
  }
  duo = du;
}

// (mass - mass0)/mass of slot i for the Bingham-erosion law (6:9451)
__device__ __forceinline__ double nsoil_of(const DevParams &P, const SortedSet &s, int i) {
  if (P.nfric != 3) return 0.0;
  const double m = s.mass[i];
  return (m - s.mass0[s.id[i]]) / m;
}

// a13. Check_Out_Domain (7:1764-1788)
__device__ __forceinline__ bool out_of_domain(const DevDem &d, double x, double y) {
  return ((x - d.xmin) * (x - d.xmax) > 0.0) || ((y - d.ymin) * (y - d.ymax) > 0.0);
}

// ---------------------------------------------------------------------------
// wave / block reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t = __shfl_xor(v, o, 64);
    v = t < v ? t : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t = __shfl_xor(v, o, 64);
    v = t > v ? t : v;
  }
  return v;
}
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Bounding box of active particles: every block of the producing kernel
// writes one partial (no atomics); k_grid reduces the partials of all blocks.
__device__ __forceinline__ void bbox_block(BBoxAcc *part, bool active, double2 p, double hs) {
  __shared__ BBoxAcc wp[GF_BLOCK / 64];
  const unsigned long long BIG = ~0ull, SMALL = 0ull;
  unsigned long long oxmin = active ? ord_of(p.x) : BIG, oxmax = active ? ord_of(p.x) : SMALL;
  unsigned long long oymin = active ? ord_of(p.y) : BIG, oymax = active ? ord_of(p.y) : SMALL;
  unsigned long long ohmax = active ? ord_of(hs) : SMALL, ohmin = active ? ord_of(hs) : BIG;
  unsigned cnt = active ? 1u : 0u;
  oxmin = wave_min_u64(oxmin);
  oxmax = wave_max_u64(oxmax);
  oymin = wave_min_u64(oymin);
  oymax = wave_max_u64(oymax);
  ohmax = wave_max_u64(ohmax);
  ohmin = wave_min_u64(ohmin);
  cnt = wave_sum_u32(cnt);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wp[wid].xmin = oxmin;
    wp[wid].xmax = oxmax;
    wp[wid].ymin = oymin;
    wp[wid].ymax = oymax;
    wp[wid].hmax = ohmax;
    wp[wid].hmin = ohmin;
    wp[wid].nactive = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BBoxAcc a = wp[0];
    for (int w = 1; w < GF_BLOCK / 64; w++) {
      a.xmin = wp[w].xmin < a.xmin ? wp[w].xmin : a.xmin;
      a.xmax = wp[w].xmax > a.xmax ? wp[w].xmax : a.xmax;
      a.ymin = wp[w].ymin < a.ymin ? wp[w].ymin : a.ymin;
      a.ymax = wp[w].ymax > a.ymax ? wp[w].ymax : a.ymax;
      a.hmax = wp[w].hmax > a.hmax ? wp[w].hmax : a.hmax;
      a.hmin = wp[w].hmin < a.hmin ? wp[w].hmin : a.hmin;
      a.nactive += wp[w].nactive;
    }
    part[blockIdx.x] = a;
  }
}

// XCD-aware block order: the dispatcher deals consecutive workgroups round
// robin to the 8 XCDs, each with its own L2.  Launched on a grid that is a
// multiple of 8, logical block (b % 8) * (grid/8) + b / 8 gives every XCD one
// contiguous range of cell-ordered particles, so neighbour gathers of
// adjacent workgroups hit the same L2.
__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// neighbour build
// ---------------------------------------------------------------------------

__global__ void k_bbox(int n, const double2 *pos, const double *hsml, const int *flag, BBoxAcc *part) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  bool act = false;
  double2 p = make_double2(0, 0);
  double hs = 0;
  if (i < n) {
    act = flag[i] == 0;
    p = pos[i];
    hs = hsml[i];
  }
  bbox_block(part, act, p, hs);
}

// Grid of the active bounding box.  Cell edge = CELL_FRAC*scale_k*max(hsml)
// *(1+1e-7), so every pair passing the predicate lies within 2 cells in x
// and y; grown if the grid would exceed cap_cells.
__device__ __forceinline__ GridInfo make_grid(const BBoxAcc &a, int scale_k, int cap_cells) {
  GridInfo g;
  if (a.nactive == 0) {
    g.xmin = 0;
    g.ymin = 0;
    g.cell = 1;
    g.hmax = 0;
    g.ncx = 1;
    g.ncy = 1;
    g.ncell = 1;
    g.uniform_h = 0;
    return g;
  }
  double xmin = val_of(a.xmin), xmax = val_of(a.xmax), ymin = val_of(a.ymin), ymax = val_of(a.ymax);
  double hmax = val_of(a.hmax);
  double cell = CELL_FRAC * scale_k * hmax * (1.0 + 1.0e-7);
  if (!(cell > 0)) cell = 1.0;
  double lx = xmax - xmin, ly = ymax - ymin;
  double ncx = floor(lx / cell) + 1, ncy = floor(ly / cell) + 1;
  if (ncx * ncy > (double)cap_cells) {
    double f = sqrt(ncx * ncy / (double)cap_cells) * 1.01;
    cell *= f;
    ncx = floor(lx / cell) + 1;
    ncy = floor(ly / cell) + 1;
  }
  g.xmin = xmin;
  g.ymin = ymin;
  g.cell = cell;
  g.hmax = hmax;
  g.ncx = (int)ncx;
  g.ncy = (int)ncy;
  g.ncell = g.ncx * g.ncy;
  g.uniform_h = a.hmin == a.hmax ? 1 : 0;
  return g;
}

// One block: reduces the bounding-box partials, stores the total and the grid.
__global__ void k_grid(int nparts, const BBoxAcc *part, BBoxAcc *acc, GridInfo *gout, int *coff, int scale_k,
                       int cap_cells) {
  __shared__ BBoxAcc lds[GF_BLOCK];
  BBoxAcc a;
  a.xmin = ~0ull;
  a.ymin = ~0ull;
  a.xmax = 0ull;
  a.ymax = 0ull;
  a.hmax = 0ull;
  a.hmin = ~0ull;
  a.nactive = 0u;
  a.pad = 0u;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    BBoxAcc b = part[k];
    a.xmin = b.xmin < a.xmin ? b.xmin : a.xmin;
    a.xmax = b.xmax > a.xmax ? b.xmax : a.xmax;
    a.ymin = b.ymin < a.ymin ? b.ymin : a.ymin;
    a.ymax = b.ymax > a.ymax ? b.ymax : a.ymax;
    a.hmax = b.hmax > a.hmax ? b.hmax : a.hmax;
    a.hmin = b.hmin < a.hmin ? b.hmin : a.hmin;
    a.nactive += b.nactive;
  }
  lds[threadIdx.x] = a;
  __syncthreads();
  for (int off = GF_BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      BBoxAcc b = lds[threadIdx.x + off];
      a.xmin = b.xmin < a.xmin ? b.xmin : a.xmin;
      a.xmax = b.xmax > a.xmax ? b.xmax : a.xmax;
      a.ymin = b.ymin < a.ymin ? b.ymin : a.ymin;
      a.ymax = b.ymax > a.ymax ? b.ymax : a.ymax;
      a.hmax = b.hmax > a.hmax ? b.hmax : a.hmax;
      a.hmin = b.hmin < a.hmin ? b.hmin : a.hmin;
      a.nactive += b.nactive;
      lds[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *acc = a;
    const GridInfo g = make_grid(a, scale_k, cap_cells);
    *gout = g;
    coff[0] = 0;
    coff[1] = g.ncell;
  }
}

// Several realisations: slot range of every realisation in the order of the
// last sort (its cells are contiguous keys), read before the grids change.
__global__ void k_rng(int R, const int *coff, const int *start, int2 *rng) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) rng[r] = make_int2(start[coff[r]], start[coff[r + 1]]);
}



// Bounding boxes of the realisations of a batch: realisation r's slot range
// (rng[r]) is cut into chunks of BOX_CHUNK slots, one work-group per (r,
// chunk) reduces its chunk to a partial (no atomics), and k_grid_real folds
// each realisation's partials into its grid.  Min / max / count do not
// depend on the order, so the grids are those of a serial reduction.
#define BOX_CHUNK 4096

__device__ __forceinline__ BBoxAcc bbox_identity() {
  BBoxAcc a;
  a.xmin = ~0ull;
  a.ymin = ~0ull;
  a.xmax = 0ull;
  a.ymax = 0ull;
  a.hmax = 0ull;
  a.hmin = ~0ull;
  a.nactive = 0u;
  a.pad = 0u;
  return a;
}

__device__ __forceinline__ void bbox_fold(BBoxAcc &a, const BBoxAcc &b) {
  a.xmin = b.xmin < a.xmin ? b.xmin : a.xmin;
  a.xmax = b.xmax > a.xmax ? b.xmax : a.xmax;
  a.ymin = b.ymin < a.ymin ? b.ymin : a.ymin;
  a.ymax = b.ymax > a.ymax ? b.ymax : a.ymax;
  a.hmax = b.hmax > a.hmax ? b.hmax : a.hmax;
  a.hmin = b.hmin < a.hmin ? b.hmin : a.hmin;
  a.nactive += b.nactive;
}

// grid: (R, nchunk) work-groups of GF_BLOCK threads; part: [R][nchunk]
__global__ void k_bbox_chunks(const int2 *rng, const double2 *pos, const double *hsml, const int *flag,
                              BBoxAcc *part) {
  __shared__ BBoxAcc wp[GF_BLOCK / 64];
  const int r = blockIdx.x, ch = blockIdx.y;
  const int2 rg = rng[r];
  const int b = rg.x + ch * BOX_CHUNK, e = min(rg.y, b + BOX_CHUNK);
  const unsigned long long BIG = ~0ull, SMALL = 0ull;
  unsigned long long oxmin = BIG, oxmax = SMALL, oymin = BIG, oymax = SMALL, ohmax = SMALL, ohmin = BIG;
  unsigned cnt = 0u;
  for (int i = b + (int)threadIdx.x; i < e; i += GF_BLOCK) {
    if (flag[i]) continue;
    const double2 p = pos[i];
    const unsigned long long ox = ord_of(p.x), oy = ord_of(p.y), oh = ord_of(hsml[i]);
    oxmin = ox < oxmin ? ox : oxmin;
    oxmax = ox > oxmax ? ox : oxmax;
    oymin = oy < oymin ? oy : oymin;
    oymax = oy > oymax ? oy : oymax;
    ohmax = oh > ohmax ? oh : ohmax;
    ohmin = oh < ohmin ? oh : ohmin;
    cnt++;
  }
  oxmin = wave_min_u64(oxmin);
  oxmax = wave_max_u64(oxmax);
  oymin = wave_min_u64(oymin);
  oymax = wave_max_u64(oymax);
  ohmax = wave_max_u64(ohmax);
  ohmin = wave_min_u64(ohmin);
  cnt = wave_sum_u32(cnt);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wp[wid].xmin = oxmin;
    wp[wid].xmax = oxmax;
    wp[wid].ymin = oymin;
    wp[wid].ymax = oymax;
    wp[wid].hmax = ohmax;
    wp[wid].hmin = ohmin;
    wp[wid].nactive = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BBoxAcc a = wp[0];
    for (int w = 1; w < GF_BLOCK / 64; w++) bbox_fold(a, wp[w]);
    part[(long)r * gridDim.y + ch] = a;
  }
}

// grid of every realisation from its partials, capped at CAP_PER_PART cells per
// particle of its slot range + 64
__global__ void k_grid_real(int R, int nchunk, const int2 *rng, const BBoxAcc *part, int scale_k, GridInfo *grid) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int2 rg = rng[r];
  BBoxAcc a = bbox_identity();
  for (int ch = 0; ch < nchunk; ch++) bbox_fold(a, part[(long)r * nchunk + ch]);
  grid[r] = make_grid(a, scale_k, CAP_PER_PART * (rg.y - rg.x) + 64);
}

__device__ __forceinline__ void cell_of(const GridInfo &g, double2 p, int &cx, int &cy) {
  cx = (int)((p.x - g.xmin) / g.cell);
  cy = (int)((p.y - g.ymin) / g.cell);
  cx = cx < 0 ? 0 : (cx >= g.ncx ? g.ncx - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= g.ncy ? g.ncy - 1 : cy);
}

// Runs of consecutive lanes of a wave that share a key: the particles arrive
// in the previous step's cell order, so the lanes of a wave mostly fall in
// a few cells.  Every lane of the wave must call these together.
struct WaveRun {
  int head;  // first lane of this lane's run
  int len;   // length of the run (valid in its head lane)
};

__device__ __forceinline__ WaveRun wave_run(long key, int lane) {
  const long prev = __shfl_up(key, 1);
  const bool is_head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(is_head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  WaveRun w;
  w.head = 63 - __clzll((long long)(heads & upto));
  const unsigned long long after = lane == 63 ? 0ull : (heads >> (lane + 1));
  const int next_head = after ? lane + __ffsll((long long)after) : 64;
  w.len = next_head - lane;
  return w;
}

// Cell keys and counts: one atomic per run of equal keys in a wave, whose
// lanes take consecutive ranks (the rank order inside a cell is fixed later
// by k_reorder).  Lanes past n carry keys of their own.
__global__ void k_keys(int n, const double2 *pos, const int *flag, const int *rid, const GridInfo *grid,
                       const int *coff, const int *ntot, int *key, int *rank, int *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)__lane_id();
  long k = -1 - lane;
  if (i < n) {
    if (flag[i]) {
      k = *ntot;
    } else {
      const int r = rid[i];
      const GridInfo g = grid[r];
      int cx, cy;
      cell_of(g, pos[i], cx, cy);
      k = coff[r] + cy * g.ncx + cx;
    }
  }
  const WaveRun w = wave_run(k, lane);
  int base = 0;
  if (lane == w.head && i < n) base = atomicAdd(&cnt[k], w.len);
  base = __shfl(base, w.head);
  if (i < n) {
    key[i] = (int)k;
    rank[i] = base + (lane - w.head);
  }
}

// exclusive scan of cnt[0..ncell] -> start[0..ncell+1]; 2048 items per block
#define SCAN_ITEMS 8
#define SCAN_TILE (GF_BLOCK * SCAN_ITEMS)

__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int &total) {
  // wave inclusive scan
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(x, o, 64);
    if (lane >= o) x += t;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < GF_BLOCK / 64; w++) {
      int t = lds[w];
      lds[w] = s;
      s += t;
    }
    lds[GF_BLOCK / 64] = s;
  }
  __syncthreads();
  int res = x - v + lds[wid];
  total = lds[GF_BLOCK / 64];
  __syncthreads();
  return res;
}

__global__ void k_scan_reduce(const int *ntot, const int *cnt, int *bsum) {
  __shared__ int lds[GF_BLOCK / 64 + 1];
  int nb = *ntot + 1;
  int base = blockIdx.x * SCAN_TILE;
  if (base >= nb) {
    if (threadIdx.x == 0) bsum[blockIdx.x] = 0;
    return;
  }
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int c = base + threadIdx.x * SCAN_ITEMS + k;
    if (c < nb) s += cnt[c];
  }
  int total;
  block_exclusive_scan(s, lds, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}


// Down-sweep: each tile's offset is the sum of the tile sums before it
// (summed by the block itself: the used tiles are few), then the tile's
// exclusive scan.  The counts are cleared as they are consumed, so the next
// build finds them zero (they start zero: cleared at allocation).
__global__ void k_scan_down(const int *ntot, int *cnt, const int *bsum, int *start, int n) {
  __shared__ int lds[GF_BLOCK / 64 + 1];
  int nb = *ntot + 1;
  int base = blockIdx.x * SCAN_TILE;
  if (base >= nb) return;
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += blockDim.x) pre += bsum[b];
  int pre_total;
  (void)block_exclusive_scan(pre, lds, pre_total);
  int v[SCAN_ITEMS];
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int c = base + threadIdx.x * SCAN_ITEMS + k;
    v[k] = (c < nb) ? cnt[c] : 0;
    s += v[k];
  }
  int total;
  int ex = block_exclusive_scan(s, lds, total) + pre_total;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int c = base + threadIdx.x * SCAN_ITEMS + k;
    if (c < nb) {
      start[c] = ex;
      cnt[c] = 0;
    }
    ex += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) start[nb] = n;
}

// exclusive scan of the realisations' cell counts -> coff[0..R]
__global__ void k_coff(int R, const GridInfo *grid, int *coff) {
  __shared__ int lds[GF_BLOCK / 64 + 1];
  int carry = 0;
  for (int base = 0; base < R; base += GF_BLOCK) {
    const int r = base + (int)threadIdx.x;
    const int v = r < R ? grid[r].ncell : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, total);
    if (r < R) coff[r] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) coff[R] = carry;
}

__global__ void k_scatter(int n, const int *key, const int *rank, const int *start, int *perm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[start[key[i]] + rank[i]] = i;
}

// Gathers the state into cell order.  perm is the counting-sort scatter
// (slot order inside a cell follows the atomics); each cell is reordered by
// previous slot -- a particle's rank among its cell's particles -- so that
// runs are deterministic (the out-of-domain bucket keeps the atomic order:
// its particles never interact).  Sources: positions / depth / smoothing
// length / pore pressure / flags from the previous step's outputs (src_new),
// mass / id / velocity from the previous sorted set (src_old).
__global__ void k_reorder(int n, const int *ntot, const int *perm, const int *key, const int *start,
                          NewSet src_new, SortedSet src_old, SortedSet dst, int ncol) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int i = perm[t];
  int c = key[i];
  int d = t;
  if (c < *ntot) {
    int b = start[c], e = start[c + 1], rank = 0;
    for (int u = b; u < e; u++) rank += perm[u] < i;
    d = b + rank;
  }
  const double2 p = src_new.pos[i];
  const double hs = src_new.hsml[i], m = src_old.mass[i];
  dst.nb[d] = make_double4(p.x, p.y, hs, m);
  dst.pos[d] = p;
  dst.vel[d] = src_old.vel[i];
  dst.h[d] = src_new.h[i];
  dst.mass[d] = m;
  dst.hsml[d] = hs;
  dst.u[d] = src_new.u[i];
  dst.uaux[d] = src_new.uaux[i];
  dst.flag[d] = src_new.flag[i];
  dst.id[d] = src_old.id[i];
  dst.rid[d] = src_old.rid[i];
  for (int k = 0; k < ncol; k++) dst.col[(long)k * n + d] = src_new.col[(long)k * n + i];
}

// ---------------------------------------------------------------------------
// neighbour iteration helper: calls f(j) for every candidate j in the 3x3
// block of cells around particle slot i (3 contiguous slot ranges).
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ void for_candidates(const GridInfo &g, const int *start, double2 p, F &&f) {
  int cx, cy;
  cell_of(g, p, cx, cy);
  int c0 = cx >= RAD ? cx - RAD : 0, c1 = cx < g.ncx - RAD ? cx + RAD : g.ncx - 1;
  int r0 = cy >= RAD ? cy - RAD : 0, r1 = cy < g.ncy - RAD ? cy + RAD : g.ncy - 1;
  for (int r = r0; r <= r1; r++) {
    int b = start[r * g.ncx + c0], e = start[r * g.ncx + c1 + 1];
    for (int j = b; j < e; j++) f(j);
  }
}

// Neighbour masks: NB_WORDS 64-bit words per particle, one bit per position
// of the particle's candidate stream (its trimmed cell-row ranges
// concatenated), set by the density pass for the candidates that pass the
// predicate.  A stream longer than 64*NB_WORDS is not masked: the passes
// re-test the predicate for it.
#define NB_WORDS NROWS

// frozen out-of-domain particle: IntForces' h_inf floor (6:6769) and
// Get_Pwp_SW's Uaux = 0 (6:11184) still apply
__device__ __forceinline__ void write_frozen(const DevParams &P, const SortedSet &s, int i, NewSet &o) {
  double hh = s.h[i];
  o.pos[i] = s.pos[i];
  o.vel[i] = s.vel[i];
  o.h[i] = hh < P.h_inf ? P.h_inf : hh;
  o.hsml[i] = s.hsml[i];
  if (P.nuaux) {
    o.uaux[i] = 0.0;
    o.u[i] = 0.0;
  } else {
    o.uaux[i] = s.uaux[i];
    o.u[i] = s.u[i];
  }
  o.flag[i] = 1;
}

// Erosion state of the terrain, per realisation r at offset r*np: the
// current Z, eroded depth and available depth (topol(1), topol(14),
// topol(15)) and the two inverse-distance sums the particles scatter each
// step (auxv(1:2, :), 6:8733-8748).
struct ErosGrid {
  double *z, *eroded, *avail;
  double *aux;  // [R][np][2]
  long np;
};

// a particle's share of the eroded depth to the 4 x 4 nodes around it
// (6:8732-8748); the sums over particles are atomic, so their order -- and
// the last bits of the node values -- follow the hardware's order
__device__ __forceinline__ void erosion_scatter(const DevDem &d, double *aux, double xx, double yy, double eb) {
  const double fx = (xx - d.xmin) / d.dx + 1, fy = (yy - d.ymin) / d.dy + 1;
  if (!(fx >= 1.0 && fx < d.nx + 1.0 && fy >= 1.0 && fy < d.ny + 1.0)) return;  // active particles are inside
  const int ipx = (int)fx, ipy = (int)fy;
  const int inx0 = ipx - 1 > 1 ? ipx - 1 : 1, iny0 = ipy - 1 > 1 ? ipy - 1 : 1;
  const int inx1 = ipx + 2 < d.nx ? ipx + 2 : d.nx, iny1 = ipy + 2 < d.ny ? ipy + 2 : d.ny;
  for (int inx = inx0; inx <= inx1; inx++)
    for (int iny = iny0; iny <= iny1; iny++) {
      const long ip = (long)(iny - 1) * d.nx + (inx - 1);
      double dx2 = xx - (d.xmin + (inx - 1) * d.dx), dy2 = yy - (d.ymin + (iny - 1) * d.dy);
      dx2 = dx2 * dx2;
      dy2 = dy2 * dy2;
      const double dist = fmax(1.e-6, sqrt(dx2 + dy2));
      unsafeAtomicAdd(aux + 2 * ip, eb / dist);
      unsafeAtomicAdd(aux + 2 * ip + 1, 1.0 / dist);
    }
}

// initial erosion state: realisation 0's Z (uploaded) to every realisation
__global__ void k_eros_init(long total, long np, const double *avail0, const double *zone_avail, const int *lawon,
                            ErosGrid g) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const long r = k / np;
  if (r > 0) g.z[k] = g.z[k - r * np];
  g.eroded[k] = 0.0;
  // const(48) or 999 with a surface law; else the terrain's erodible depths
  g.avail[k] = (lawon[r] || !zone_avail) ? avail0[r] : zone_avail[k - r * np];
  g.aux[2 * k] = 0.0;
  g.aux[2 * k + 1] = 0.0;
}

// the terrain update after the particles (6:8790-8806); clears the sums
__global__ void k_erode_nodes(long total, ErosGrid g) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const double2 a = reinterpret_cast<const double2 *>(g.aux)[k];
  if (a.x == 0.0 && a.y == 0.0) return;
  if (a.y > 1.e-6) {
    double xnum = a.x / a.y;
    const double av = g.avail[k];
    if (xnum > av) xnum = av;
    if (av > 0.) {
      g.eroded[k] = g.eroded[k] + xnum;
      g.z[k] = g.z[k] - xnum;
      g.avail[k] = av - xnum;
    }
  }
  reinterpret_cast<double2 *>(g.aux)[k] = make_double2(0.0, 0.0);
}

// After the stages: final update (7:908-922), final depth (7:926-931), domain
// flag, erosion (ExtForces_SW(1), 6:8724-8761), pore pressure npwp = 1
// (6:11176-11310), u clamp (7:971-980), h_upgrade sle = 1 (7:2925-2931).
// Writes the particle's outputs.  EROS: (exl, eyl) is the last stage's
// external force, (inx, iny) the pressure force.
template <bool EROS = false>
__device__ __forceinline__ void finish_particle(const DevParams &P, const DevDem &dem, const SortedSet &s, int i,
                                                double Si, double rfi, double2 x, double2 v, double u, double Dx,
                                                double Dy, double Dvx, double Dvy, double Du, bool flagged,
                                                NewSet &o, bool &active_out, double2 &xo, double &hso,
                                                double exl = 0.0, double eyl = 0.0, double inx = 0.0,
                                                double iny = 0.0, double *eaux = nullptr) {
  const double dt = P.dt;
  const double2 x0 = s.pos[i], v0 = s.vel[i];
  const double u0 = s.u[i], h0 = s.h[i], uaux1 = s.uaux[i], mi = s.mass[i], hi = s.hsml[i];
  double hnew, mnew = mi;
  if (!flagged) {
    x.x = x0.x + P.fin * Dx;
    x.y = x0.y + P.fin * Dy;
    v.x = v0.x + P.fin * Dvx;
    v.y = v0.y + P.fin * Dvy;
    u = u0 + P.fin * Du;
    hnew = Si;
  } else {
    hnew = rfi;  // frozen mid-step: keeps its clamped stage depth
  }
  if (out_of_domain(dem, x.x, x.y)) flagged = true;
  const double hS = hnew;  // Drho_RK is taken before ExtForces_SW(1) (7:928)
  if (EROS && P.eros_on && !flagged) {
    double er = 0.0, e1, e2, e3;
    ext_force<true>(P, dem, x.x, x.y, v.x, v.y, hnew, u, uaux1, inx, iny, 0.0, nsoil_of(P, s, i), e1, e2, e3, exl,
                    eyl, true, &er);
    erosion_scatter(dem, eaux, x.x, x.y, dt * er);
    mnew = mi + dt * er * mi / hnew;
    if (P.eros_single) hnew = hnew + dt * er;
    s.mass[i] = mnew;
    s.nb[i].w = mnew;  // XSPH reads the masses after erosion (7:987)
  }
  double uaux_new = uaux1;
  if (P.npwp) {
    // the column is updated by k_pwp_column after the stages (it needs the
    // step's depth change); the bed value and u follow it there
    o.dh[i] = hS - h0;
    if (flagged) uaux_new = 0.0;
    u = uaux_new;
  } else if (P.nuaux) {
    if (flagged) {
      uaux_new = 0.0;
    } else {
      double hh = hnew, dhh = hS - h0, hinv;
      if (hh < P.comp) {
        hh = P.comp;
        hinv = 1.0 / P.comp;
      } else {
        hinv = 1 / hh;
      }
      double betaw = P.Bfact * hinv * hinv / P.hrelpw2;
      double e = exp(-betaw * dt);
      uaux_new = uaux1 * e + (dhh / dt) * (1. - e) / betaw;
      if (uaux_new > hh) uaux_new = hh;
      if (uaux_new < hh * P.alphaS * P.fpwp_min) uaux_new = hh * P.alphaS * P.fpwp_min;
    }
    u = uaux_new;
  } else if (!flagged) {
    if (u > hnew) u = hnew;
    if (u < 0.0) u = 0.0;
  }
  double hs = hi;
  if (P.sle == 1 && !flagged) hs = 2.0 * pow(mnew / hnew, 1. / 2);
  o.pos[i] = x;
  o.vel[i] = v;
  o.h[i] = hnew;
  o.hsml[i] = hs;
  o.u[i] = u;
  o.uaux[i] = uaux_new;
  o.flag[i] = flagged ? 1 : 0;
  active_out = !flagged;
  xo = x;
  hso = hs;
}

// stage update of position, velocity and pore pressure (7:811-833)
__device__ __forceinline__ void stage_update(const DevParams &P, int st, double2 x0, double2 v0, double u0, double h0,
                                             double rfi, double ex, double ey, double du, double inx, double iny,
                                             double2 &x, double2 &v, double &u) {
  const double f1 = P.f1[st], dt = P.dt;
  x.x = x0.x + f1 * dt * v.x;
  x.y = x0.y + f1 * dt * v.y;
  v.x = v0.x + f1 * dt * (ex + inx);
  v.y = v0.y + f1 * dt * (ey + iny);
  u = u0 + f1 * dt * du;
  if (P.nuaux == 0) {
    double rr = (st == 0) ? h0 : rfi;
    if (u > rr) u = rr;
    else if (u < 0.0) u = 0.0;
  }
}

// ---------------------------------------------------------------------------
// a6..a10, a12, a13: the RK stages per particle (every basal law except
// Bingham, whose stages couple particles): DEM lookup, gravity, rheology,
// limiter, stage update, then the end-of-step updates
// ---------------------------------------------------------------------------
// Pr holds one parameter block per realisation; the lanes of a wave are
// served one realisation at a time (normally one pass), so every parameter
// read is a scalar load.  part = nullptr: no bounding-box partials (several
// realisations: k_bbox_chunks computes their boxes).
#ifndef RK_WAVES_PER_SIMD
#define RK_WAVES_PER_SIMD 1
#endif
// waves per SIMD the specialised kernel is compiled for (its registers fit
// three waves with a few spilled words; measured faster than two)
#ifndef RK5_WAVES_PER_SIMD
#define RK5_WAVES_PER_SIMD 3
#endif
template <bool EROS, int LAW>
__global__ __launch_bounds__(GF_BLOCK, LAW == LAW_FRIC5 ? RK5_WAVES_PER_SIMD : RK_WAVES_PER_SIMD) void k_rk(int n, const DevParams *Pr, DevDem dem, SortedSet s,
                                                  const double *S, const double2 *ain, NewSet o, BBoxAcc *part,
                                                  const int *stopr, ErosGrid eg) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  bool active_out = false;
  double2 xo = make_double2(0, 0);
  double hso = 0;
  const int myr = i < n ? s.rid[i] : 0;
  for (bool todo = i < n; todo;) {
    const int r0 = __builtin_amdgcn_readfirstlane(myr);
    if (myr != r0) continue;
    todo = false;
    const DevParams &P = Pr[r0];
    if (stopr && stopr[r0]) {  // a realisation whose case has stopped does not change
      o.pos[i] = s.pos[i];
      o.vel[i] = s.vel[i];
      o.h[i] = s.h[i];
      o.hsml[i] = s.hsml[i];
      o.u[i] = s.u[i];
      o.uaux[i] = s.uaux[i];
      o.flag[i] = s.flag[i];
      active_out = !s.flag[i];
      xo = s.pos[i];
      hso = s.hsml[i];
    } else if (s.flag[i]) {
      write_frozen(P, s, i, o);
    } else {
      const double2 in = ain[i];
      const double inx = in.x, iny = in.y;
      const double Si = S[i];
      const double rfi = Si < P.h_inf ? P.h_inf : Si;
      const double2 x0 = s.pos[i], v0 = s.vel[i];
      const double u0 = s.u[i], h0 = s.h[i], uaux1 = s.uaux[i];
      double2 x = x0, v = v0;
      double u = u0, ex = 0.0, ey = 0.0, du = 0.0;
      double Dx = 0, Dy = 0, Dvx = 0, Dvy = 0, Du = 0;
      bool flagged = false;
      double tau_unused = 0.0;
      const double nsoil = LAW == LAW_FRIC5 ? 0.0 : nsoil_of(P, s, i);
      for (int st = 0; st < P.nstages; st++) {
        stage_update(P, st, x0, v0, u0, h0, rfi, ex, ey, du, inx, iny, x, v, u);
        if (out_of_domain(dem, x.x, x.y)) {
          flagged = true;
          break;
        }
        ext_force<EROS, LAW>(P, dem, x.x, x.y, v.x, v.y, rfi, u, uaux1, inx, iny, tau_unused, nsoil, ex, ey, du);
        const double f2 = P.f2[st];
        Dvx += f2 * (ex + inx);
        Dvy += f2 * (ey + iny);
        Dx += f2 * v.x;
        Dy += f2 * v.y;
        Du += f2 * du;
      }
      finish_particle<EROS>(P, dem, s, i, Si, rfi, x, v, u, Dx, Dy, Dvx, Dvy, Du, flagged, o, active_out, xo, hso,
                            ex, ey, inx, iny, EROS ? eg.aux + 2 * eg.np * r0 : nullptr);
    }
  }
  if (part) bbox_block(part, active_out, xo, hso);
}

// ---------------------------------------------------------------------------
// Bingham (nfric = 4): ExtForces_SW applies cf4b = Tau0/dens after every
// branch (6:9619-9622) but sets Tau0 only in the intermediate regime
// (6:9616), so a particle in the viscous or cohesive regime uses the Tau0 of
// the last particle, in particle index order, that was intermediate in the
// same ExtForces call -- or in an earlier call.  Each stage therefore runs as
//   k_bingham_stage(phase 0): stage update, regime, candidate Tau0 per id
//   last-valid scan over particle ids (with the carry of earlier calls)
//   k_bingham_stage(phase 1): basal friction with the resolved Tau0
// ---------------------------------------------------------------------------
struct BinghamState {
  double2 *in;   // pressure force (stage invariant)
  double2 *x, *v, *ex;
  double *u, *du;
  double *D;     // [n][5] Dx Dy Dvx Dvy Du
  int *fl;       // flagged in mid-step
  double *cand;  // [n] by caller id: Tau0 or NaN
  double *tau;   // [n] by caller id: resolved Tau0
  double *carry; // [R]: each realisation's Tau0 carried from its earlier calls
  double *bsum;  // per scan tile: last valid value (NaN if none)
  int4 *tiles;   // [ntiles]: realisation, first id, end id of each scan tile
  int *rtile;    // [R + 1]: first tile of each realisation
  int ntiles;
};

__global__ __launch_bounds__(GF_BLOCK) void k_bingham_init(int n, SortedSet s, const double2 *ain, BinghamState b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b.in[i] = ain[i];
  b.x[i] = s.pos[i];
  b.v[i] = s.vel[i];
  b.ex[i] = make_double2(0.0, 0.0);
  b.u[i] = s.u[i];
  b.du[i] = 0.0;
  for (int k = 0; k < 5; k++) b.D[5 * i + k] = 0.0;
  b.fl[i] = 0;
}

// Pr: the realisations' parameter blocks on the device (dt_sph and the RK
// weight change with the adaptive time step between steps); each wave runs
// its realisations one after the other (their slot ranges are contiguous)
__global__ __launch_bounds__(GF_BLOCK) void k_bingham_stage(int n, int phase, int st, const DevParams *Pr,
                                                             DevDem dem, SortedSet s, const double *S,
                                                             BinghamState b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int myr = i < n ? s.rid[i] : 0;
  for (bool todo = i < n; todo;) {
    const int r0 = __builtin_amdgcn_readfirstlane(myr);
    if (myr != r0) continue;
    todo = false;
    const DevParams &P = Pr[r0];
    const int id = s.id[i];
    if (s.flag[i] || b.fl[i]) {
      if (phase == 0) b.cand[id] = __longlong_as_double(0x7ff8000000000000ll);
      continue;
    }
    const double Si = S[i];
    const double rfi = Si < P.h_inf ? P.h_inf : Si;
    const double2 in = b.in[i];
    if (phase == 0) {
      double2 x = b.x[i], v = b.v[i], ex = b.ex[i];
      double u = b.u[i];
      stage_update(P, st, s.pos[i], s.vel[i], s.u[i], s.h[i], rfi, ex.x, ex.y, b.du[i], in.x, in.y, x, v, u);
      b.x[i] = x;
      b.v[i] = v;
      b.u[i] = u;
      if (out_of_domain(dem, x.x, x.y)) {
        b.fl[i] = 1;
        b.cand[id] = __longlong_as_double(0x7ff8000000000000ll);
        continue;
      }
      b.cand[id] = bingham_candidate(P, rfi, v.x, v.y);
    } else {
      double2 x = b.x[i], v = b.v[i];
      double u = b.u[i], ex, ey, du;
      double tau = b.tau[id];
      ext_force(P, dem, x.x, x.y, v.x, v.y, rfi, u, s.uaux[i], in.x, in.y, tau, 0.0, ex, ey, du);
      b.u[i] = u;
      b.ex[i] = make_double2(ex, ey);
      b.du[i] = du;
      const double f2 = P.f2[st];
      double *D = b.D + 5 * i;
      D[2] += f2 * (ex + in.x);
      D[3] += f2 * (ey + in.y);
      D[0] += f2 * v.x;
      D[1] += f2 * v.y;
      D[4] += f2 * du;
    }
  }
}

__global__ __launch_bounds__(GF_BLOCK) void k_bingham_finish(int n, const DevParams *Pr, DevDem dem, SortedSet s,
                                                              const double *S, BinghamState b, NewSet o,
                                                              BBoxAcc *part) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  bool active_out = false;
  double2 xo = make_double2(0, 0);
  double hso = 0;
  const int myr = i < n ? s.rid[i] : 0;
  for (bool todo = i < n; todo;) {
    const int r0 = __builtin_amdgcn_readfirstlane(myr);
    if (myr != r0) continue;
    todo = false;
    const DevParams &P = Pr[r0];
    if (s.flag[i]) {
      write_frozen(P, s, i, o);
    } else {
      const double Si = S[i];
      const double rfi = Si < P.h_inf ? P.h_inf : Si;
      const double *D = b.D + 5 * i;
      finish_particle(P, dem, s, i, Si, rfi, b.x[i], b.v[i], b.u[i], D[0], D[1], D[2], D[3], D[4], b.fl[i] != 0,
                      o, active_out, xo, hso);
    }
  }
  if (part) bbox_block(part, active_out, xo, hso);
}

// last-valid inclusive scan over particle ids: tau[k] = cand of the largest
// k' <= k with cand[k'] a number, else the carry of earlier calls
#define LV_ITEMS 8
#define LV_TILE (GF_BLOCK * LV_ITEMS)

__device__ __forceinline__ bool is_valid(double v) { return v == v; }

// Scan tiles never cross a realisation: each realisation carries its own
// Tau0 (realisations never interact), and ids are stacked realisation after
// realisation.  Tile b covers ids tiles[b].y .. tiles[b].z - 1 of
// realisation tiles[b].x.
__global__ void k_lv_reduce(const int4 *tiles, const double *cand, double *bsum) {
  __shared__ double lds[GF_BLOCK];
  const int4 tl = tiles[blockIdx.x];
  double last = __longlong_as_double(0x7ff8000000000000ll);
  for (int k = 0; k < LV_ITEMS; k++) {
    int idx = tl.y + threadIdx.x * LV_ITEMS + k;
    if (idx < tl.z && is_valid(cand[idx])) last = cand[idx];
  }
  lds[threadIdx.x] = last;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = __longlong_as_double(0x7ff8000000000000ll);
    for (int t = 0; t < GF_BLOCK; t++)
      if (is_valid(lds[t])) r = lds[t];
    bsum[blockIdx.x] = r;
  }
}

// one thread per realisation, sequential over its tiles: bsum becomes the
// value valid before each tile; the realisation's carry becomes the last
// valid value of this call
__global__ void k_lv_top(int R, const int *rtile, double *bsum, double *carry) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double c = carry[r];
  for (int b = rtile[r]; b < rtile[r + 1]; b++) {
    double t = bsum[b];
    bsum[b] = c;
    if (is_valid(t)) c = t;
  }
  carry[r] = c;
}

__global__ void k_lv_down(const int4 *tiles, const double *cand, const double *bsum, double *tau) {
  __shared__ double lds[GF_BLOCK];
  const int4 tl = tiles[blockIdx.x];
  double last = __longlong_as_double(0x7ff8000000000000ll);
  for (int k = 0; k < LV_ITEMS; k++) {
    int idx = tl.y + threadIdx.x * LV_ITEMS + k;
    if (idx < tl.z && is_valid(cand[idx])) last = cand[idx];
  }
  lds[threadIdx.x] = last;
  __syncthreads();
  // Hillis-Steele inclusive scan of "last valid" over the block's threads
  for (int off = 1; off < GF_BLOCK; off <<= 1) {
    double mine = lds[threadIdx.x];
    double other = threadIdx.x >= off ? lds[threadIdx.x - off] : __longlong_as_double(0x7ff8000000000000ll);
    __syncthreads();
    if (!is_valid(mine)) lds[threadIdx.x] = other;
    __syncthreads();
  }
  double c = threadIdx.x > 0 ? lds[threadIdx.x - 1] : __longlong_as_double(0x7ff8000000000000ll);
  if (!is_valid(c)) c = bsum[blockIdx.x];
  for (int k = 0; k < LV_ITEMS; k++) {
    int idx = tl.y + threadIdx.x * LV_ITEMS + k;
    if (idx < tl.z) {
      if (is_valid(cand[idx])) c = cand[idx];
      tau[idx] = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Neighbour gathers.  A particle's candidates lie in the 5 cell rows
// cy-2 .. cy+2; in each row they are one contiguous slot range of the
// cell-sorted set, trimmed to the columns the support circle reaches in that
// row (x within +-w of the particle, w^2 = T^2 - dy^2 with dy the distance to
// the row band and T = scale_k*(h_i + h_max)/2 an upper bound of every pair's
// support radius: candidates outside cannot be neighbours).  The 5 ranges,
// concatenated in row order, form the particle's candidate stream; stream
// position p of row rr is slot p + delta[rr].  The density pass tests every
// candidate and records the neighbours as bits of the stream; the force and
// XSPH passes walk the set bits.  One stream per particle (rather than one
// loop per row) keeps the lanes of a wave busy: the wave waits for its
// longest stream, not for the longest of each row.  Loads are issued in
// batches of NB_BATCH independent candidates before any of them is used, the
// batch bookkeeping is branch-free, and every sum runs in stream order (rows
// in order, slots ascending within a row).
// ---------------------------------------------------------------------------
#ifndef NB_BATCH
#define NB_BATCH 4
#endif
#define GATHER_MAX_BLOCK 512
// minimum waves per SIMD of each gather (register budget), if set
#ifdef DENSITY_WAVES
#define DENSITY_LB __launch_bounds__(GATHER_MAX_BLOCK, DENSITY_WAVES)
#else
#define DENSITY_LB __launch_bounds__(GATHER_MAX_BLOCK)
#endif
// the force pass fits five waves per SIMD without spilling
#ifndef PFORCE_WAVES
#define PFORCE_WAVES 5
#endif
#define PFORCE_LB __launch_bounds__(GATHER_MAX_BLOCK, PFORCE_WAVES)
#ifdef XSPH_WAVES
#define XSPH_LB __launch_bounds__(GATHER_MAX_BLOCK, XSPH_WAVES)
#else
#define XSPH_LB __launch_bounds__(GATHER_MAX_BLOCK)
#endif

// Every gather splits its sum into two halves, each summed in stream order,
// and the result is half 0 + half 1.  The density pass (which walks stream
// positions) puts the 4-position batches 0, 2, 4, ... in half 0 and 1, 3,
// 5, ... in half 1; the force and XSPH passes (which walk the neighbour
// mask) put the neighbours of even ordinal among the set bits of their mask
// word in half 0 and the others in half 1.  A context runs the halves
// either on two adjacent lanes (TM = 2: more loads in flight when the
// particles alone fill each SIMD only a few waves deep) or both on one lane
// (TM = 1: large batches); both give the same bits, for every work-group
// size.
#define SPLIT 2
// particles from which one lane per particle is the default
#ifndef TEAM1_MIN_N
#define TEAM1_MIN_N 1000000
#endif
// default work-group size of the gathers (gfsph_autotune may change it)
#ifndef GATHER_BLOCK
#define GATHER_BLOCK 64
#endif

// particle and lane in its team of this thread
template <int TM>
__device__ __forceinline__ void team_index(int &i, int &t) {
  const int g = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  i = g / TM;
  t = g % TM;
}

// half 0 + half 1: the lane's two partial sums, or the pair's (lane 0 adds
// its own and its partner's; lane 1 gets the same value in the other order)
template <int TM>
__device__ __forceinline__ double halves_sum(const double (&a)[SPLIT / TM]) {
  if constexpr (TM == 1)
    return a[0] + a[1];
  else
    return a[0] + __shfl_xor(a[0], 1);
}

template <int TM>
__device__ __forceinline__ unsigned long long halves_or(unsigned long long v) {
  if constexpr (TM == 1) {
    return v;
  } else {
    const unsigned lo = __shfl_xor((unsigned)v, 1), hi = __shfl_xor((unsigned)(v >> 32), 1);
    return v | ((unsigned long long)hi << 32) | lo;
  }
}

// The neighbours of mask word v of ordinal parity h among its set bits (h = 0:
// the first, third, ... set bit): v masked by the inclusive prefix parity of
// its bits (or its complement).
__device__ __forceinline__ unsigned long long ordinal_half(unsigned long long v, int h) {
  unsigned long long x = v;
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return v & (h ? ~x : x);
}

struct CandStream {
  int off[NROWS];  // first stream position of each row (off[0] = 0)
  int inc[NROWS];  // slot of stream position p: p + sum of inc[rr] over the rows with off[rr] <= p
  int total;       // length of the stream
};

__device__ __forceinline__ CandStream cand_stream(const GridInfo &g, const int *start, double2 p, double hi,
                                                  int scale_k) {
  int cx, cy;
  cell_of(g, p, cx, cy);
  const double T = scale_k * (hi + g.hmax) * 0.5 * (1.0 + 1.0e-9);
  const double T2 = T * T;
  const int cmin = cx >= RAD ? cx - RAD : 0, cmax = cx < g.ncx - RAD ? cx + RAD : g.ncx - 1;
  CandStream cs;
  int tot = 0, dprev = 0;
#pragma unroll
  for (int rr = 0; rr < NROWS; rr++) {
    int bb = 0, cnt = 0;
    const int r = cy - RAD + rr;
    if (r >= 0 && r < g.ncy) {
      const double y0 = g.ymin + r * g.cell, y1 = y0 + g.cell;
      const double dy = p.y < y0 ? y0 - p.y : (p.y > y1 ? p.y - y1 : 0.0);
      const double w2 = T2 - dy * dy;
      if (w2 > 0.0) {
        const double w = sqrt(w2);
        int c0 = (int)floor((p.x - w - g.xmin) / g.cell), c1 = (int)floor((p.x + w - g.xmin) / g.cell);
        c0 = c0 < cmin ? cmin : c0;
        c1 = c1 > cmax ? cmax : c1;
        if (c1 >= c0) {
          bb = start[r * g.ncx + c0];
          cnt = start[r * g.ncx + c1 + 1] - bb;
        }
      }
    }
    cs.off[rr] = tot;
    cs.inc[rr] = bb - tot - dprev;  // increments of delta[rr] = bb - off[rr]
    dprev = bb - tot;
    tot += cnt;
  }
  cs.total = tot;
  return cs;
}

// a stream longer than the mask holds: the passes re-test the predicate
__device__ __forceinline__ bool stream_full(const CandStream &cs) { return cs.total > 64 * NB_WORDS; }

// The density pass stores each particle's stream (32 bytes: the row
// increments and the 16-bit row offsets and length) for the force and XSPH
// passes of the step; a full stream stores a marker and is recomputed.
static_assert(NROWS == 5, "the stream record holds 5 rows");
#define STREAM_FULL_MARK 0xffffu

__device__ __forceinline__ void store_stream(int4 *rec, int i, const CandStream &cs) {
  const bool full = stream_full(cs);
  const unsigned o12 = (unsigned)cs.off[1] | ((unsigned)cs.off[2] << 16);
  const unsigned o34 = (unsigned)cs.off[3] | ((unsigned)cs.off[4] << 16);
  const unsigned tot = full ? STREAM_FULL_MARK : (unsigned)cs.total;
  rec[2 * i] = make_int4(cs.inc[0], cs.inc[1], cs.inc[2], cs.inc[3]);
  rec[2 * i + 1] = make_int4(cs.inc[4], (int)o12, (int)o34, (int)tot);
}

// false: a full stream (the caller recomputes it)
__device__ __forceinline__ bool load_stream(const int4 *rec, int i, CandStream &cs) {
  const int4 a = rec[2 * i], b = rec[2 * i + 1];
  if ((unsigned)b.w == STREAM_FULL_MARK) return false;
  cs.inc[0] = a.x;
  cs.inc[1] = a.y;
  cs.inc[2] = a.z;
  cs.inc[3] = a.w;
  cs.inc[4] = b.x;
  cs.off[0] = 0;
  cs.off[1] = (int)((unsigned)b.y & 0xffffu);
  cs.off[2] = (int)((unsigned)b.y >> 16);
  cs.off[3] = (int)((unsigned)b.z & 0xffffu);
  cs.off[4] = (int)((unsigned)b.z >> 16);
  cs.total = b.w;
  return true;
}

// Slot of stream position p: p + delta of the last row with off[rr] <= p
// (an empty row shares its off with the next one, so its increment is
// cancelled by the next row's).  Summed as masked increments so that it
// stays in registers (a select between two of the row values would be
// turned into an indexed load).
__device__ __forceinline__ int slot_of(const CandStream &cs, int p) {
  int d = p + cs.inc[0];
#pragma unroll
  for (int rr = 1; rr < NROWS; rr++) d += (p >= cs.off[rr] ? -1 : 0) & cs.inc[rr];
  return d;
}

// Separation and mean smoothing length of a pair, rounded as in pair_pred
// (dr = dx*dx + dy*dy without contraction).
__device__ __forceinline__ void pair_geom(double2 pi_, double hi, const double4 &q, double &dx, double &dy,
                                          double &dr, double &mh) {
  dx = __dsub_rn(pi_.x, q.x);
  dy = __dsub_rn(pi_.y, q.y);
  dr = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
  mh = __dadd_rn(hi, q.z) / 2.;
}

// Thresholds of the square-root-free predicate for support radius t: with
// lo = t*t*(1-2^-46) and hi = t*t*(1+2^-46) (each rounded), dr < lo implies
// sqrt(dr) < t and dr > hi implies sqrt(dr) > t after every rounding
// involved; only dr inside [lo, hi] needs the exact test of pair_pred.
struct PredBand {
  double t, lo, hi;
};

__device__ __forceinline__ PredBand pred_band(double t) {
  const double tt = __dmul_rn(t, t);
  PredBand pb;
  pb.t = t;
  pb.lo = __dmul_rn(tt, 1.0 - 0x1p-46);
  pb.hi = __dmul_rn(tt, 1.0 + 0x1p-46);
  return pb;
}

// The mask words of a particle's stream (words past the stream read as 0).
__device__ __forceinline__ void load_words(const CandStream &cs, const unsigned long long *mrow,
                                           unsigned long long (&w)[NB_WORDS], unsigned long long lane_bits) {
  const int nw = (cs.total + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NB_WORDS; k++) w[k] = k < nw ? (mrow[k] & lane_bits) : 0ull;
}

// Calls f(js, ok) for batches of NB_BATCH neighbours taken from the mask
// words w of a particle in stream order; js are slots.  Unused entries of a
// batch have ok = false and js = self (a valid index), so f may read them
// unconditionally.  A batch takes its neighbours from one word (an emptied
// word ends the batch and is shifted out before the next), so entry u of a
// batch has the ordinal parity u & 1 among its word's set bits.
template <typename F>
__device__ __forceinline__ void word_batches(const CandStream &cs, unsigned long long (&w)[NB_WORDS], int self,
                                             F &&f) {
  const int nw = (cs.total + 63) >> 6;
  int k = 0;
  for (;;) {
    while (w[0] == 0ull && k + 1 < nw) {
#pragma unroll
      for (int m = 0; m + 1 < NB_WORDS; m++) w[m] = w[m + 1];
      w[NB_WORDS - 1] = 0ull;
      k++;
    }
    if (w[0] == 0ull) break;
    const int pbase = k << 6;
    int js[NB_BATCH];
    bool ok[NB_BATCH];
#pragma unroll
    for (int u = 0; u < NB_BATCH; u++) {
      ok[u] = w[0] != 0ull;
      const int sl = slot_of(cs, pbase + __ffsll((long long)w[0]) - 1);
      js[u] = ok[u] ? sl : self;
      w[0] &= w[0] - 1ull;
    }
    f(js, ok);
  }
}

// Every neighbour in the stream, with the exact predicate, for particles
// whose streams are too long for the mask (not batched: rare).
// Positions t, t+T, t+2T, ... of the stream (T = 1: all of them).
template <int T = SPLIT, typename F>
__device__ __forceinline__ void stream_dense(const CandStream &cs, int scale_k, int i, int t, double2 pi_, double hi,
                                             const double4 *nb, F &&f) {
  for (int p = t; p < cs.total; p += T) {
    const int j = slot_of(cs, p);
    if (j == i) continue;
    const double4 q = nb[j];
    double dx, dy, d, mh;
    if (pair_pred(scale_k, pi_, make_double2(q.x, q.y), hi, q.z, dx, dy, d, mh))
      f(j, q, dx, dy, __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), mh);
  }
}

// r = sqrt(dr) for the pair sums: the hardware reciprocal-root estimate
// (relative error below 2^-24), one Newton-Raphson step and one correction
// of the root.  An emulation in exact arithmetic over 20,000 arguments with
// estimates at that error bound gave the correctly rounded root every time.
// dr = 0 (a particle and itself, or two coincident particles) is taken as
// 2^-1000: its root 2^-500 gives the kernel's q = 0 values exactly.
__device__ __forceinline__ double root_pair(double dr) {
  dr = fmax(dr, 0x1p-1000);
  double y = __builtin_amdgcn_rsq(dr);
  const double hx = 0.5 * dr;
  const double e = fma(-(hx * y), y, 0.5);
  y = fma(y, e, y);
  double g = dr * y;
  const double d = fma(-g, g, dr);
  return fma(d, 0.5 * y, g);
}

// Kernel value alone (the density and XSPH sums) for a compile-time kernel
// type, from the squared separation dr: the polynomials of kernel_TOOLS
// ("5a My_SPH_Tools_2019.f90":2062-2180) in nested form, one branch-free
// select between the two pieces of the cubic spline.  Agrees with
// kernel_TOOLS to a few ulps for q = r/h <= 2, the only values the sums use
// (a neighbour has r < 2h); beyond q = 2 the result is not the kernel.
template <int SKF, bool UH = false>
__device__ __forceinline__ double sph_w_t(double kc, double pi, double dr, double hsml, double ihu = 0.0) {
  if (SKF == 1 || SKF == 4) {
    const double r = root_pair(dr);
    const double ih = UH ? ihu : rcp_nr(hsml);
    const double q = r * ih;
    const double factor = kc * (ih * ih);
    if (SKF == 1) {
      const double w1 = fma(q * q, fma(0.5, q, -1.0), 2.0 / 3.0);  // 2/3 - q^2 + q^3/2
      const double t = 2.0 - q;
      const double w2 = (t * t) * (t * (1.0 / 6.0));                // (2 - q)^3 / 6
      return factor * (q <= 1.0 ? w1 : w2);
    } else {
      const double b = fma(-0.5, q, 1.0);                            // 1 - q/2
      const double b2 = b * b;
      return factor * ((b2 * b2) * fma(2.0, q, 1.0));                // b^4 (2q + 1)
    }
  } else {
    double w, gx, gy;
    sph_kernel(SKF, pi, sqrt(dr), 0.0, 0.0, hsml, w, gx, gy);
    return w;
  }
}

// Kernel gradient alone for a pair sum, rounded as sph_kernel_t rounds it: the cubic spline's inner piece needs no 1/r, the
// outer one takes 1/r from two Newton-Raphson steps on the estimate.
template <int SKF, bool UH = false>
__device__ __forceinline__ void sph_grad_nb(double kc, double pi, double dr, double dx, double dy, double hsml,
                                            double ihu, double &gx, double &gy) {
  if (SKF == 1) {
    dr = fmax(dr, 0x1p-1000);
    double y = __builtin_amdgcn_rsq(dr);
    const double hx = 0.5 * dr;
    double e = fma(-(hx * y), y, 0.5);
    y = fma(y, e, y);
    e = fma(-(hx * y), y, 0.5);
    y = fma(y, e, y);
    double r = dr * y;
    const double d = fma(-r, r, dr);
    r = fma(d, 0.5 * y, r);
    const double ih = UH ? ihu : rcp_nr(hsml);
    const double ih2 = ih * ih;
    const double q = r * ih;
    const double factor = kc * ih2;
    const double a1 = factor * (-2. + 1.5 * q) * ih2;
    const double t = 2. - q;
    const double a2 = -(factor * 0.5 * (t * t) * ih) * y;
    const double a = q <= 1.0 ? a1 : (q <= 2.0 ? a2 : 0.0);
    gx = a * dx;
    gy = a * dy;
  } else {
    double w;
    sph_kernel_t<SKF, UH>(kc, pi, dr, dx, dy, hsml, w, gx, gy, ihu);
  }
}

// One batch of NB_BATCH consecutive stream positions k0 .. k0+3 of the
// density pass: every candidate takes the predicate and the kernel value,
// the neighbours' m_j W_ij are summed into acc in stream order, and the
// returned nibble has bit u set for a neighbour at position k0 + u.
// UNIFORM: all smoothing lengths are equal, so the support radius
// scale_k*(hi+hj)/2 = scale_k*hi and the kernel's 1/h are the particle's own.
template <int SKF, bool UNIFORM>
__device__ __forceinline__ unsigned density_batch(const CandStream &cs, const DevParams &P, int self, int k0,
                                                  double2 pi_, double hi, const PredBand &pu, double ihi,
                                                  const double4 *nb, double &acc) {
  int j[NB_BATCH];
  double4 q[NB_BATCH];
#pragma unroll
  for (int u = 0; u < NB_BATCH; u++) {
    j[u] = slot_of(cs, min(k0 + u, cs.total - 1));
    q[u] = nb[j[u]];
  }
  double dx[NB_BATCH], dy[NB_BATCH], dr[NB_BATCH], mh[NB_BATCH];
  PredBand pb[NB_BATCH];
  bool in[NB_BATCH], band[NB_BATCH];
  bool anyband = false;
#pragma unroll
  for (int u = 0; u < NB_BATCH; u++) {
    pair_geom(pi_, hi, q[u], dx[u], dy[u], dr[u], mh[u]);
    pb[u] = UNIFORM ? pu : pred_band(P.scale_k * mh[u]);
    in[u] = dr[u] < pb[u].lo;
    band[u] = !in[u] && !(dr[u] > pb[u].hi);
    anyband = anyband || band[u];
  }
  if (__any(anyband)) {
#pragma unroll
    for (int u = 0; u < NB_BATCH; u++)
      if (band[u]) in[u] = sqrt(dr[u]) < pb[u].t;
  }
  unsigned nib = 0u;
#pragma unroll
  for (int u = 0; u < NB_BATCH; u++) {
    const bool nbr = in[u] && k0 + u < cs.total && j[u] != self;
    nib |= nbr ? (1u << u) : 0u;
    const double t = acc + q[u].w * sph_w_t<SKF, UNIFORM>(P.kc, P.pi, dr[u], mh[u], ihi);
    acc = nbr ? t : acc;
  }
  return nib;
}

// Density sum over the stream with the neighbour-mask build.  The stream is
// cut into batches of 4 positions; batch b belongs to half b & 1, and each
// half is summed in stream order (acc[0], acc[1]).  TM = 1: one lane takes
// the batches two at a time, the first into acc[0], the second into acc[1];
// TM = 2: lane h takes the batches of half h into acc[0].  Mask word k holds
// stream positions 64k .. 64k+63; its bits are gathered in a register while
// the word's positions are walked.
template <int SKF, bool UNIFORM, int TM>
__device__ __forceinline__ void density_stream(const CandStream &cs, const DevParams &P, int self, int h,
                                               double2 pi_, double hi, const double4 *nb,
                                               unsigned long long (&wd)[NB_WORDS], double (&acc)[2]) {
  const PredBand pu = pred_band(P.scale_k * hi);
  const double ihi = rcp_nr(hi);
  acc[0] = 0.0;
  acc[1] = 0.0;
#pragma unroll
  for (int wk = 0; wk < NB_WORDS; wk++) {
    unsigned long long bits = 0ull;
    for (int k = (TM == 1 ? 0 : NB_BATCH * h); k < 64 && 64 * wk + k < cs.total; k += SPLIT * NB_BATCH) {
      const int k0 = 64 * wk + k;
      unsigned nib = density_batch<SKF, UNIFORM>(cs, P, self, k0, pi_, hi, pu, ihi, nb, acc[0]);
      if constexpr (TM == 1)
        nib |= density_batch<SKF, UNIFORM>(cs, P, self, k0 + NB_BATCH, pi_, hi, pu, ihi, nb, acc[1]) << NB_BATCH;
      bits |= (unsigned long long)nib << k;
    }
    wd[wk] = bits;
  }
}

// a4 density with the neighbour-mask build (pair predicate of 5a:1010-1019)
template <int SKF, int TM>
__global__ DENSITY_LB void k_density_m(int n, DevParams P, Cells cl, SortedSet s, double *S,
                                                                 double *Pout, unsigned long long *mask,
                                                                 int4 *srec, double4 *nbp) {
  int i, t;
  team_index<TM>(i, t);
  if (i >= n) return;
  if (s.flag[i]) {
    if (t == 0) {
      S[i] = s.h[i];
      Pout[i] = 0.0;
    }
    return;
  }
  const int rr_ = s.rid[i];
  const GridInfo g = cl.grid[rr_];
  const int *start = cl.start + cl.coff[rr_];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  const CandStream cs = cand_stream(g, start, pi_, hi, P.scale_k);
  if (t == 0) store_stream(srec, i, cs);
  double a[SPLIT / TM];
  if (stream_full(cs)) {
#pragma unroll
    for (int hh = 0; hh < SPLIT / TM; hh++) {
      double acc = 0.0;
      stream_dense(cs, P.scale_k, i, t + hh * TM, pi_, hi, s.nb,
                   [&](int j, const double4 &q, double dx, double dy, double dr, double mh) {
                     acc = acc + q.w * sph_w_t<SKF>(P.kc, P.pi, dr, mh);
                   });
      a[hh] = acc;
    }
  } else {
    unsigned long long wd[NB_WORDS];
#pragma unroll
    for (int k = 0; k < NB_WORDS; k++) wd[k] = 0ull;
    double two[2];
    if (g.uniform_h)
      density_stream<SKF, true, TM>(cs, P, i, t, pi_, hi, s.nb, wd, two);
    else
      density_stream<SKF, false, TM>(cs, P, i, t, pi_, hi, s.nb, wd, two);
#pragma unroll
    for (int hh = 0; hh < SPLIT / TM; hh++) a[hh] = two[hh];
    unsigned long long *mrow = mask + NB_WORDS * (long)i;
    const int nw = (cs.total + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NB_WORDS; k++) {
      const unsigned long long v = halves_or<TM>(wd[k]);
      if (k < nw && k % TM == t) mrow[k] = v;
    }
  }
  const double acc = halves_sum<TM>(a);
  if (t != 0) return;
  const double Si = sph_self_w(SKF, P.pi, hi) * s.mass[i] + acc;
  S[i] = Si;
  const double rf = Si < P.h_inf ? P.h_inf : Si;
  const double p = P.cgra05 * rf * rf;
  const double Pi = (P.pa_sph == 2) ? p / (rf * rf) : p;
  Pout[i] = Pi;
  nbp[i] = make_double4(pi_.x, pi_.y, Pi, s.mass[i]);
}

// pressure-force contribution of neighbour j (IntForces_SW pair term)
template <int PA>
__device__ __forceinline__ void pforce_term(double Pi, double rfi, double Pj, double Sj, double h_inf, double mj,
                                            double gx, double gy, double &inx, double &iny) {
  if (PA == 2) {
    const double c = -(Pi + Pj);
    inx = inx + mj * (c * gx);
    iny = iny + mj * (c * gy);
  } else if (PA == 1) {
    const double rfj = Sj < h_inf ? h_inf : Sj;
    const double rhoij = 1.e0 / (rfi * rfj);
    inx = inx + mj * ((-(Pi + Pj) * gx) * rhoij);
    iny = iny + mj * ((-(Pi + Pj) * gy) * rhoij);
  } else {  // pa_sph == 4
    const double rfj = Sj < h_inf ? h_inf : Sj;
    inx = inx - ((Pj - Pi) * gx * mj / rfj);
    iny = iny - ((Pj - Pi) * gy * mj / rfj);
  }
}

// a5 pressure force (IntForces_SW pair sum, 6:6892-6941) -> ain[i]
template <int SKF, int PA, int TM>
__global__ PFORCE_LB void k_pforce_m(int n, DevParams P, Cells cl,
                                                                const unsigned long long *mask, SortedSet s,
                                                                const double *S, const double *Pp, double2 *ain,
                                                                const int4 *srec, const double4 *nbp) {
  int i, t;
  team_index<TM>(i, t);
  if (i >= n) return;
  if (s.flag[i]) {
    if (t == 0) ain[i] = make_double2(0.0, 0.0);
    return;
  }
  const int rr_ = s.rid[i];
  const GridInfo g = cl.grid[rr_];
  const int *start = cl.start + cl.coff[rr_];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  CandStream cs;
  if (!load_stream(srec, i, cs)) cs = cand_stream(g, start, pi_, hi, P.scale_k);
  const double Si = S[i];
  const double rfi = Si < P.h_inf ? P.h_inf : Si;
  const double Pi = Pp[i];
  double ax[SPLIT / TM], ay[SPLIT / TM];
#pragma unroll
  for (int hh = 0; hh < SPLIT / TM; hh++) ax[hh] = ay[hh] = 0.0;
  if (stream_full(cs)) {
#pragma unroll
    for (int hh = 0; hh < SPLIT / TM; hh++)
      stream_dense(cs, P.scale_k, i, t + hh * TM, pi_, hi, s.nb,
                   [&](int j, const double4 &q, double dx, double dy, double dr, double mh) {
                     double w, gx, gy;
                     sph_kernel_t<SKF>(P.kc, P.pi, dr, dx, dy, mh, w, gx, gy);
                     pforce_term<PA>(Pi, rfi, Pp[j], PA == 2 ? 0.0 : S[j], P.h_inf, q.w, gx, gy, ax[hh], ay[hh]);
                   });
  } else {
    unsigned long long w[NB_WORDS];
    load_words(cs, mask + NB_WORDS * (long)i, w, ~0ull);
    if constexpr (TM == 2) {
#pragma unroll
      for (int k = 0; k < NB_WORDS; k++) w[k] = ordinal_half(w[k], t);
    }
    const double ihi = rcp_nr(hi);
    // an unused entry (ok = false: the particle itself) adds m = 0 times a
    // finite term, which leaves the sum's bits as they are
    auto body = [&](auto uh, const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
      constexpr bool UH = decltype(uh)::value;
      double4 q[NB_BATCH];
      double pj[NB_BATCH], sj[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        q[u] = s.nb[js[u]];
        pj[u] = Pp[js[u]];
        sj[u] = PA == 2 ? 0.0 : S[js[u]];
      }
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        double dx, dy, dr, mh, gx, gy;
        pair_geom(pi_, hi, q[u], dx, dy, dr, mh);
        sph_grad_nb<SKF, UH>(P.kc, P.pi, dr, dx, dy, mh, ihi, gx, gy);
        const int hh = TM == 1 ? (u & 1) : 0;
        pforce_term<PA>(Pi, rfi, pj[u], sj[u], P.h_inf, ok[u] ? q[u].w : 0.0, gx, gy, ax[hh], ay[hh]);
      }
    };
    // equal smoothing lengths and pa_sph = 2: a neighbour's position, pressure
    // and mass come as one 32-byte record (written by the density pass)
    auto body_p = [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
      double4 r[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) r[u] = nbp[js[u]];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        double dx, dy, dr, mh, gx, gy;
        pair_geom(pi_, hi, make_double4(r[u].x, r[u].y, hi, r[u].w), dx, dy, dr, mh);
        sph_grad_nb<SKF, true>(P.kc, P.pi, dr, dx, dy, mh, ihi, gx, gy);
        const int hh = TM == 1 ? (u & 1) : 0;
        pforce_term<PA>(Pi, rfi, r[u].z, 0.0, P.h_inf, ok[u] ? r[u].w : 0.0, gx, gy, ax[hh], ay[hh]);
      }
    };
    if (PA == 2 && nbp && g.uniform_h)
      word_batches(cs, w, i, body_p);
    else if (g.uniform_h)
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::true_type{}, js, ok);
      });
    else
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::false_type{}, js, ok);
      });
  }
  const double inx = halves_sum<TM>(ax), iny = halves_sum<TM>(ay);
  if (t == 0) ain[i] = make_double2(inx, iny);
}

// a11 XSPH (av_vel, 7:3042-3099) with the frozen step-start geometry
template <int SKF, int TM>
__global__ XSPH_LB void k_xsph_m(int n, DevParams P, Cells cl,
                                                              const unsigned long long *mask, SortedSet s, NewSet o,
                                                              double2 *vout, const int *stopr, const int4 *srec) {
  int i, t;
  team_index<TM>(i, t);
  if (i >= n) return;
  double2 vi = o.vel[i];
  if (s.flag[i] || o.flag[i] || s.id[i] < P.npoin0 || (stopr && stopr[s.rid[i]])) {
    if (t == 0) vout[i] = vi;
    return;
  }
  const int rr_ = s.rid[i];
  const GridInfo g = cl.grid[rr_];
  const int *start = cl.start + cl.coff[rr_];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  CandStream cs;
  if (!load_stream(srec, i, cs)) cs = cand_stream(g, start, pi_, hi, P.scale_k);
  const double rhoi = o.h[i];
  const bool skip0 = P.npoin0 > 0;
  double hx[SPLIT / TM], hy[SPLIT / TM];
#pragma unroll
  for (int hh = 0; hh < SPLIT / TM; hh++) hx[hh] = hy[hh] = 0.0;
  if (stream_full(cs)) {
#pragma unroll
    for (int hh = 0; hh < SPLIT / TM; hh++)
      stream_dense(cs, P.scale_k, i, t + hh * TM, pi_, hi, s.nb,
                   [&](int j, const double4 &q, double dx, double dy, double dr, double mh) {
                     if (skip0 && s.id[j] < P.npoin0) return;
                     const double w = sph_w_t<SKF>(P.kc, P.pi, dr, mh);
                     const double2 vj = o.vel[j];
                     const double c = (2 * q.w) * w / (rhoi + o.h[j]);
                     hx[hh] = hx[hh] - c * (vi.x - vj.x);
                     hy[hh] = hy[hh] - c * (vi.y - vj.y);
                   });
  } else {
    unsigned long long w[NB_WORDS];
    load_words(cs, mask + NB_WORDS * (long)i, w, ~0ull);
    if constexpr (TM == 2) {
#pragma unroll
      for (int k = 0; k < NB_WORDS; k++) w[k] = ordinal_half(w[k], t);
    }
    const double ihi = rcp_nr(hi);
    // an unused entry adds c = 0 times a finite difference (bits unchanged)
    auto body = [&](auto uh, const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
      constexpr bool UH = decltype(uh)::value;
      double4 q[NB_BATCH];
      double2 vj[NB_BATCH];
      double hj[NB_BATCH];
      bool use[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        q[u] = s.nb[js[u]];
        vj[u] = o.vel[js[u]];
        hj[u] = o.h[js[u]];
        use[u] = ok[u] && !(skip0 && s.id[js[u]] < P.npoin0);
      }
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        double dx, dy, dr, mh;
        pair_geom(pi_, hi, q[u], dx, dy, dr, mh);
        const double w = sph_w_t<SKF, UH>(P.kc, P.pi, dr, mh, ihi);
        const double mj = use[u] ? q[u].w : 0.0;
        const double c = (2 * mj) * w * rcp_nr(rhoi + hj[u]);
        const int hh = TM == 1 ? (u & 1) : 0;
        hx[hh] = hx[hh] - c * (vi.x - vj[u].x);
        hy[hh] = hy[hh] - c * (vi.y - vj[u].y);
      }
    };
    if (g.uniform_h)
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::true_type{}, js, ok);
      });
    else
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::false_type{}, js, ok);
      });
  }
  const double ax = halves_sum<TM>(hx), ay = halves_sum<TM>(hy);
  if (t != 0) return;
  vi.x = vi.x + 0.3 * ax;
  vi.y = vi.y + 0.3 * ay;
  vout[i] = vi;
}

// a12 h_upgrade (7:2866-2961) for sle = 2 and 3 over the step-start pair set
// (the neighbour masks), with the step's final velocities, depths and masses
// (after erosion), before XSPH; writes the new smoothing length of every
// particle inside the domain.
//   sle = 2: dh/dt = -(h/2)(1/h_d) dh_d/dt, i.e. vcc_i = sum_j m_j (v_j - v_i)
//            . grad_i W_ij / h_j and hsml += dt (hsml/2) vcc_i, undone if it
//            would reach 0 (7:2893-2919);
//   sle = 3: McDougall & Hungr, hsml = 4 / sqrt(S_i / N_i) with N_i the
//            particle's neighbours and S_i its own h_i/m_i added N_i times
//            (7:2933-2959).  A particle without neighbours keeps its
//            smoothing length (the reference takes 0/0 there and carries a
//            NaN smoothing length into the next neighbour search).
// The sums run in two halves (ordinal parity of a neighbour among its mask
// word's set bits) on one lane, as the XSPH pass runs them.
template <int SKF>
__global__ XSPH_LB void k_hsml_m(int n, const DevParams *Pr, Cells cl, const unsigned long long *mask, SortedSet s,
                                 NewSet o, const int *stopr, const int4 *srec) {
  const int i = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (i >= n) return;
  const int rr_ = s.rid[i];
  if (s.flag[i] || o.flag[i] || (stopr && stopr[rr_])) return;
  const DevParams &P = Pr[rr_];
  const GridInfo g = cl.grid[rr_];
  const int *start = cl.start + cl.coff[rr_];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  CandStream cs;
  if (!load_stream(srec, i, cs)) cs = cand_stream(g, start, pi_, hi, P.scale_k);
  if (P.sle == 3) {
    int N = 0;
    if (stream_full(cs)) {
      stream_dense<1>(cs, P.scale_k, i, 0, pi_, hi, s.nb,
                      [&](int, const double4 &, double, double, double, double) { N++; });
    } else {
      const unsigned long long *mrow = mask + NB_WORDS * (long)i;
      const int nw = (cs.total + 63) >> 6;
#pragma unroll
      for (int k = 0; k < NB_WORDS; k++)
        if (k < nw) N += __popcll(mrow[k]);
    }
    if (N == 0) return;
    const double ci = o.h[i] / s.mass[i];
    double S = 0.0;
    for (int k = 0; k < N; k++) S = S + ci;
    const double Ls = sqrt(S / N);
    o.hsml[i] = 4 / Ls;
    return;
  }
  const double2 vi = o.vel[i];
  double a[SPLIT] = {0.0, 0.0};
  auto term = [&](double mj, double gx, double gy, double2 vj, double hj) {
    double hv = (vj.x - vi.x) * gx;
    hv = hv + (vj.y - vi.y) * gy;
    return mj * hv / hj;
  };
  if (stream_full(cs)) {
#pragma unroll
    for (int hh = 0; hh < SPLIT; hh++)
      stream_dense(cs, P.scale_k, i, hh, pi_, hi, s.nb,
                   [&](int j, const double4 &q, double dx, double dy, double dr, double mh) {
                     double w, gx, gy;
                     sph_kernel_t<SKF>(P.kc, P.pi, dr, dx, dy, mh, w, gx, gy);
                     a[hh] = a[hh] + term(q.w, gx, gy, o.vel[j], o.h[j]);
                   });
  } else {
    unsigned long long w[NB_WORDS];
    load_words(cs, mask + NB_WORDS * (long)i, w, ~0ull);
    const double ihi = rcp_nr(hi);
    // an unused entry (the particle itself) adds 0 * (v_i - v_i) . grad / h_i
    auto body = [&](auto uh, const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
      constexpr bool UH = decltype(uh)::value;
      double4 q[NB_BATCH];
      double2 vj[NB_BATCH];
      double hj[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        q[u] = s.nb[js[u]];
        vj[u] = o.vel[js[u]];
        hj[u] = o.h[js[u]];
      }
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        double dx, dy, dr, mh, gx, gy;
        pair_geom(pi_, hi, q[u], dx, dy, dr, mh);
        sph_grad_nb<SKF, UH>(P.kc, P.pi, dr, dx, dy, mh, ihi, gx, gy);
        a[u & 1] = a[u & 1] + term(ok[u] ? q[u].w : 0.0, gx, gy, vj[u], hj[u]);
      }
    };
    if (g.uniform_h)
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::true_type{}, js, ok);
      });
    else
      word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
        body(std::false_type{}, js, ok);
      });
  }
  const double vcc = a[0] + a[1];
  const double dh = (hi / 2) * vcc;
  double hs = hi + P.dt * dh;
  if (hs <= 0) hs = hs - P.dt * dh;
  o.hsml[i] = hs;
}

// ---------------------------------------------------------------------------
// Pressure re-evaluated in every RK stage (ic_K0 = 1 or ic_cosTheta = 1).
// RK45 calls IntForces_SW in every stage (7:849-866) with the pair geometry
// of the step start; its pressures then change with the stage: g along the
// terrain normal at the stage position (6:6774-6783) and the K0 factor by the
// sign of the stage velocity divergence (6:6736-6760, 6:6857-6863).  Each
// stage is
//   k_stg_update   stage update, domain flag, the particle's pressure
//   k_stg_k0       (ic_K0) velocity divergence over the neighbours, K0
//   k_pforce_m     the pressure force with these pressures
//   k_stg_ext      external forces and the RK sums
// and k_stg_finish ends the step as k_rk does.  A particle that leaves the
// domain keeps the pressure of its last stage inside (p is not recomputed
// for it), as its neighbours see it in the reference.  Buffers: BinghamState.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GF_BLOCK) void k_stg_init(int n, SortedSet s, BinghamState b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b.x[i] = s.pos[i];
  b.v[i] = s.vel[i];
  b.ex[i] = make_double2(0.0, 0.0);
  b.u[i] = s.u[i];
  b.du[i] = 0.0;
  for (int k = 0; k < 5; k++) b.D[5 * i + k] = 0.0;
  b.fl[i] = 0;
}

__global__ __launch_bounds__(GF_BLOCK) void k_stg_update(int n, int st, const DevParams *Pr, DevDem dem,
                                                          SortedSet s, const double *S, const double2 *ain,
                                                          BinghamState b, double *Pp, const int *stopr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = s.rid[i];
  if (s.flag[i] || b.fl[i] || (stopr && stopr[r])) return;
  const DevParams &P = Pr[r];
  const double Si = S[i];
  const double rfi = Si < P.h_inf ? P.h_inf : Si;
  // the first stage's weight is 0: its force terms are not read
  const double2 in = st == 0 ? make_double2(0.0, 0.0) : ain[i];
  double2 x = b.x[i], v = b.v[i];
  const double2 ex = b.ex[i];
  double u = b.u[i];
  stage_update(P, st, s.pos[i], s.vel[i], s.u[i], s.h[i], rfi, ex.x, ex.y, b.du[i], in.x, in.y, x, v, u);
  b.x[i] = x;
  b.v[i] = v;
  b.u[i] = u;
  if (out_of_domain(dem, x.x, x.y)) {
    b.fl[i] = 1;
    return;
  }
  double cgra05 = P.cgra05;
  if (P.costheta) {
    TopoPt tp;
    get_z_topo(dem, x.x, x.y, false, tp);
    const double deng = sqrt(1. + tp.Zx * tp.Zx + tp.Zy * tp.Zy);
    cgra05 = 0.5 * (P.cgra / deng);
  }
  const double p = cgra05 * rfi * rfi;
  if (P.ic_k0)
    b.cand[i] = p;  // the K0 factor follows in k_stg_k0
  else
    Pp[i] = (P.pa_sph == 2) ? p / (rfi * rfi) : p;
}

template <int SKF>
__global__ __launch_bounds__(GF_BLOCK) void k_stg_k0(int n, const DevParams *Pr, Cells cl,
                                                      const unsigned long long *mask, SortedSet s, const double *S,
                                                      BinghamState b, double *Pp, const int *stopr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = s.rid[i];
  if (s.flag[i] || b.fl[i] || (stopr && stopr[r])) return;
  const DevParams &P = Pr[r];
  const GridInfo g = cl.grid[r];
  const int *start = cl.start + cl.coff[r];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  const CandStream cs = cand_stream(g, start, pi_, hi, P.scale_k);
  const double2 vi = b.v[i];
  double dv = 0.0;  // sum_j m_j (v_i - v_j) . grad_i W_ij
  auto term = [&](double mj, double2 vj, double gx, double gy) {
    double vcc = (vi.x - vj.x) * gx;
    vcc = vcc + (vi.y - vj.y) * gy;
    dv = dv + mj * vcc;
  };
  if (stream_full(cs)) {
    stream_dense<1>(cs, P.scale_k, i, 0, pi_, hi, s.nb,
                    [&](int j, const double4 &q, double dx, double dy, double dr, double mh) {
                      double w, gx, gy;
                      sph_kernel_t<SKF>(P.kc, P.pi, dr, dx, dy, mh, w, gx, gy);
                      term(q.w, b.v[j], gx, gy);
                    });
  } else {
    unsigned long long w[NB_WORDS];
    load_words(cs, mask + NB_WORDS * (long)i, w, ~0ull);
    word_batches(cs, w, i, [&](const int(&js)[NB_BATCH], const bool(&ok)[NB_BATCH]) {
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        if (!ok[u]) continue;
        const double4 q = s.nb[js[u]];
        double dx, dy, dr, mh, wk, gx, gy;
        pair_geom(pi_, hi, q, dx, dy, dr, mh);
        sph_kernel_t<SKF>(P.kc, P.pi, dr, dx, dy, mh, wk, gx, gy);
        term(q.w, b.v[js[u]], gx, gy);
      }
    });
  }
  const double Si = S[i];
  const double rfi = Si < P.h_inf ? P.h_inf : Si;
  const double p = b.cand[i] * (dv >= 0 ? P.k0a : P.k0p);
  Pp[i] = (P.pa_sph == 2) ? p / (rfi * rfi) : p;
}

template <bool EROS>
__global__ __launch_bounds__(GF_BLOCK) void k_stg_ext(int n, int st, const DevParams *Pr, DevDem dem, SortedSet s,
                                                       const double *S, const double2 *ain, BinghamState b,
                                                       const int *stopr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = s.rid[i];
  if (s.flag[i] || b.fl[i] || (stopr && stopr[r])) return;
  const DevParams &P = Pr[r];
  const double Si = S[i];
  const double rfi = Si < P.h_inf ? P.h_inf : Si;
  const double2 in = ain[i], x = b.x[i], v = b.v[i];
  double u = b.u[i], ex, ey, du;
  ext_force<EROS>(P, dem, x.x, x.y, v.x, v.y, rfi, u, s.uaux[i], in.x, in.y, 0.0, nsoil_of(P, s, i), ex, ey, du);
  b.u[i] = u;
  b.ex[i] = make_double2(ex, ey);
  b.du[i] = du;
  const double f2 = P.f2[st];
  double *D = b.D + 5 * i;
  D[2] += f2 * (ex + in.x);
  D[3] += f2 * (ey + in.y);
  D[0] += f2 * v.x;
  D[1] += f2 * v.y;
  D[4] += f2 * du;
}

template <bool EROS>
__global__ __launch_bounds__(GF_BLOCK) void k_stg_finish(int n, const DevParams *Pr, DevDem dem, SortedSet s,
                                                          const double *S, const double2 *ain, BinghamState b,
                                                          NewSet o, BBoxAcc *part, const int *stopr, ErosGrid eg) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  bool active_out = false;
  double2 xo = make_double2(0, 0);
  double hso = 0;
  if (i < n) {
    const int r = s.rid[i];
    const DevParams &P = Pr[r];
    if (stopr && stopr[r]) {  // a realisation whose case has stopped does not change
      o.pos[i] = s.pos[i];
      o.vel[i] = s.vel[i];
      o.h[i] = s.h[i];
      o.hsml[i] = s.hsml[i];
      o.u[i] = s.u[i];
      o.uaux[i] = s.uaux[i];
      o.flag[i] = s.flag[i];
      active_out = !s.flag[i];
      xo = s.pos[i];
      hso = s.hsml[i];
    } else if (s.flag[i]) {
      write_frozen(P, s, i, o);
    } else {
      const double Si = S[i];
      const double rfi = Si < P.h_inf ? P.h_inf : Si;
      const double *D = b.D + 5 * i;
      const double2 ex = b.ex[i], in = ain[i];
      finish_particle<EROS>(P, dem, s, i, Si, rfi, b.x[i], b.v[i], b.u[i], D[0], D[1], D[2], D[3], D[4],
                            b.fl[i] != 0, o, active_out, xo, hso, ex.x, ex.y, in.x, in.y,
                            EROS ? eg.aux + 2 * eg.np * r : nullptr);
    }
  }
  if (part) bbox_block(part, active_out, xo, hso);
}

// ---------------------------------------------------------------------------
// Adjoint (reverse mode) of one step, for the frictional / Voellmy laws
// (nfric 5, 6) with pa_sph = 2, skf = 1, icpwp = 0, sle = 0.  With the
// neighbour set of the step fixed, one step is
//   S     = W(0) m_i + sum_j m_j W_ij                      (density)
//   a_i   = -g sum_j m_j grad W_ij                         (pressure force, P = g/2)
//   (x', v')_i = stages(x_i, v_i, max(S_i, h_inf), a_i; mu, xi)   (per particle)
//   v''_i = v'_i + 0.3 sum_j 2 m_j W_ij / (S_i + S_j) (v'_j - v'_i)  (XSPH)
// and its vector-Jacobian product runs backwards through these four maps:
// the pair maps by gathers over the same neighbour masks (each particle
// collects both sides of its pairs, so no atomics), the per-particle map by
// forward-mode dual numbers over its 9 inputs.
// ---------------------------------------------------------------------------
#define AD_N 9  // x0 (2), v0 (2), rf, a (2), mu, xi

struct Dual {
  double v;
  double d[AD_N];
};

__device__ __forceinline__ Dual dconst(double v) {
  Dual r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = 0.0;
  return r;
}
__device__ __forceinline__ Dual dvar(double v, int k) {
  Dual r = dconst(v);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ Dual operator+(const Dual &a, const Dual &b) {
  Dual r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual &a, const Dual &b) {
  Dual r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator-(const Dual &a) {
  Dual r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = -a.d[k];
  return r;
}
__device__ __forceinline__ Dual operator*(const Dual &a, const Dual &b) {
  Dual r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator*(double a, const Dual &b) {
  Dual r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a * b.d[k];
  return r;
}
__device__ __forceinline__ Dual operator*(const Dual &b, double a) { return a * b; }
__device__ __forceinline__ Dual operator/(const Dual &a, const Dual &b) {
  Dual r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
__device__ __forceinline__ Dual operator/(const Dual &a, double b) {
  Dual r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a.d[k] / b;
  return r;
}
__device__ __forceinline__ Dual operator/(double a, const Dual &b) { return dconst(a) / b; }
__device__ __forceinline__ Dual operator+(const Dual &a, double b) { return a + dconst(b); }
__device__ __forceinline__ Dual operator+(double b, const Dual &a) { return a + dconst(b); }
__device__ __forceinline__ Dual operator-(const Dual &a, double b) { return a - dconst(b); }
__device__ __forceinline__ Dual operator-(double b, const Dual &a) { return dconst(b) - a; }
__device__ __forceinline__ Dual dsqrt(const Dual &a) {
  Dual r;
  r.v = sqrt(a.v);
  const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int k = 0; k < AD_N; k++) r.d[k] = a.d[k] * h;
  return r;
}
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(const Dual &a) { return a.v; }
__device__ __forceinline__ double tsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ Dual tsqrt(const Dual &a) { return dsqrt(a); }
template <typename T>
__device__ __forceinline__ T tconst(double v);
template <>
__device__ __forceinline__ double tconst<double>(double v) {
  return v;
}
template <>
__device__ __forceinline__ Dual tconst<Dual>(double v) {
  return dconst(v);
}

// bilinear nodal slopes of Get_Z_Topo (same arithmetic as get_z_topo)
template <typename T>
__device__ __forceinline__ void slope_at(const DevDem &d, const T &xx, const T &yy, T &Zx, T &Zy) {
  int ipx = (int)((val(xx) - d.xmin) / d.dx + 1);
  int ipy = (int)((val(yy) - d.ymin) / d.dy + 1);
  if (ipx == d.nx && fabs(d.xmax - val(xx)) <= 1.e-3) ipx = ipx - 1;
  if (ipy == d.ny && fabs(d.ymax - val(yy)) <= 1.e-3) ipy = ipy - 1;
  ipx = ipx < 1 ? 1 : (ipx > d.nx - 1 ? d.nx - 1 : ipx);
  ipy = ipy < 1 ? 1 : (ipy > d.ny - 1 ? d.ny - 1 : ipy);
  long n1 = (long)(ipy - 1) * d.nx + (ipx - 1);
  long n2 = n1 + 1, n4 = n1 + d.nx, n3 = n4 + 1;
  double xref = d.xmin + (ipx - 1) * d.dx;
  double yref = d.ymin + (ipy - 1) * d.dy;
  T xi = (xx - xref) / d.dx, eta = (yy - yref) / d.dy;
  T sh1 = (1. - xi) * (1. - eta), sh2 = xi * (1. - eta), sh3 = xi * eta, sh4 = (1. - xi) * eta;
  double2 a = d.zxy[n1], b = d.zxy[n2], c = d.zxy[n3], e = d.zxy[n4];
  Zx = sh1 * a.x + sh2 * b.x + sh3 * c.x + sh4 * e.x;
  Zy = sh1 * a.y + sh2 * b.y + sh3 * c.y + sh4 * e.y;
}

// ext_force for nfric 5/6 without curvature or pore pressure, with mu
// (const(9)) and xi (const(3)) as arguments; T = double reproduces ext_force.
template <typename T>
__device__ __forceinline__ void ext_frictional(const DevParams &P, const DevDem &dem, const T &xx, const T &yy,
                                               const T &vx1, const T &vy1, const T &rho_i, const T &inx,
                                               const T &iny, const T &mu, const T &xi_voellmy, T &exo, T &eyo) {
  const double dt = P.dt;
  T h, hu, hv, hinv, vmod, hvmod, hvmodinv;
  const double gravz = P.cgra;
  T frictx = tconst<T>(0), fricty = tconst<T>(0), cf5 = tconst<T>(0);
  h = rho_i;
  hu = h * vx1;
  hv = h * vy1;
  if (val(h) < P.comp) {
    h = tconst<T>(P.comp);
    hu = tconst<T>(0);
    hv = tconst<T>(0);
    hinv = tconst<T>(0);
    vmod = tconst<T>(0);
    hvmod = tconst<T>(0);
    hvmodinv = tconst<T>(0);
  } else {
    hinv = 1 / h;
    hvmod = tsqrt(hu * hu + hv * hv);
    vmod = hvmod * hinv;
    hvmodinv = (val(hvmod) >= 1.e-5) ? 1. / hvmod : tconst<T>(0);
  }
  T Zx, Zy;
  slope_at(dem, xx, yy, Zx, Zy);
  T exx = tconst<T>(0) + (-P.cgra) * Zx;
  T exy = tconst<T>(0) + (-P.cgra) * Zy;
  T gtanfi = (gravz + 0.0) * mu * P.fact_dens;
  if (val(gtanfi) < 0.0) gtanfi = tconst<T>(0);
  cf5 = gtanfi;
  frictx = cf5 * hu * hvmodinv * h;
  fricty = cf5 * hv * hvmodinv * h;
  if (val(xi_voellmy) > 0.0) {
    T vm2 = vx1 * vx1 + vy1 * vy1;
    vmod = (val(vm2) > 0.0) ? tsqrt(vm2) : tconst<T>(0);
    frictx = frictx + P.cgra * vmod * vx1 / xi_voellmy;
    fricty = fricty + P.cgra * vmod * vy1 / xi_voellmy;
  }
  T dcx = -frictx * hinv, dcy = -fricty * hinv;
  T ax = exx + inx, ay = exy + iny;
  if (val(vmod) > 1.e-5) {
    T vsx = vx1 + dt * (ax + dcx), vsy = vy1 + dt * (ay + dcy);
    T vdot = vsx * vx1 + vsy * vy1;
    if (val(vdot) < 0) {
      T tx = vx1 / vmod, ty = vy1 / vmod;
      T accelT = ax * tx + ay * ty;
      T deccelT = dcx * tx + dcy * ty;
      if (val(accelT) * val(deccelT) < 0) {
        dcx = -vx1 / dt - ax;
        dcy = -vy1 / dt - ay;
        frictx = -dcx * h;
        fricty = -dcy * h;
      }
    }
  } else {
    T accelT = tsqrt(ax * ax + ay * ay);
    if (val(accelT) > 1.e-5) {
      T ctf = cf5;
      frictx = ctf * ax / accelT;
      fricty = ctf * ay / accelT;
      T fricT = tsqrt(frictx * frictx + fricty * fricty);
      if (val(accelT) <= val(fricT)) {
        frictx = ax;
        fricty = ay;
      }
    } else {
      frictx = tconst<T>(0);
      fricty = tconst<T>(0);
    }
    frictx = frictx * h;
    fricty = fricty * h;
  }
  exo = exx - frictx / h;
  eyo = exy - fricty / h;
}

// all RK stages of one particle and the final update (k_rk for these laws);
// mid = true if the particle left the domain during the stages
template <typename T>
__device__ __forceinline__ void stage_map(const DevParams &P, const DevDem &dem, const T &x0x, const T &x0y,
                                          const T &v0x, const T &v0y, const T &rf, const T &inx, const T &iny,
                                          const T &mu, const T &xi, T &xox, T &xoy, T &vox, T &voy, bool &mid) {
  const double dt = P.dt;
  T xx = x0x, xy = x0y, vx = v0x, vy = v0y;
  T ex = tconst<T>(0), ey = tconst<T>(0);
  T Dx = tconst<T>(0), Dy = tconst<T>(0), Dvx = tconst<T>(0), Dvy = tconst<T>(0);
  mid = false;
  for (int st = 0; st < P.nstages; st++) {
    const double f1 = P.f1[st];
    xx = x0x + f1 * dt * vx;
    xy = x0y + f1 * dt * vy;
    vx = v0x + f1 * dt * (ex + inx);
    vy = v0y + f1 * dt * (ey + iny);
    if (out_of_domain(dem, val(xx), val(xy))) {
      mid = true;
      break;
    }
    ext_frictional(P, dem, xx, xy, vx, vy, rf, inx, iny, mu, xi, ex, ey);
    const double f2 = P.f2[st];
    Dvx = Dvx + f2 * (ex + inx);
    Dvy = Dvy + f2 * (ey + iny);
    Dx = Dx + f2 * vx;
    Dy = Dy + f2 * vy;
  }
  xox = x0x + P.fin * Dx;
  xoy = x0y + P.fin * Dy;
  vox = v0x + P.fin * Dvx;
  voy = v0y + P.fin * Dvy;
}

// Bingham (nfric 4) in T, the same arithmetic as bingham_tau0,
// bingham_candidate and ext_force's Bingham branch: the intermediate-regime
// Tau0 from (tauy0, K) = (const(6), const(7)), the law's friction, the
// limiter (at rest its friction bound is |friction| itself).
template <typename T>
__device__ __forceinline__ bool bingham_tau0_t(const T &tauy0, const T &K, const T &hvmod, const T &hfrict,
                                               T &Tau0) {
  const T a_hat = ((6. * K) * hvmod) / (tauy0 * (hfrict * hfrict));
  if (val(a_hat) > 50 || val(a_hat) < 0.01) return false;
  const double a2 = 48., c2 = 65.;
  const T b2 = (-32.) * a_hat - 114.;
  const T xh = (-b2 - tsqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
  Tau0 = tauy0 / xh;
  return true;
}

template <typename T>
__device__ __forceinline__ T tmax_const(const T &a, double b) {
  return val(a) >= b ? a : tconst<T>(b);
}

template <typename T>
__device__ __forceinline__ bool bingham_candidate_t(const DevParams &P, const T &rho_i, const T &vx1, const T &vy1,
                                                    const T &tauy0, const T &K, T &cand) {
  T h = rho_i, hvmod;
  if (val(h) < P.comp) {
    h = tconst<T>(P.comp);
    hvmod = tconst<T>(0);
  } else {
    const T hu = h * vx1, hv = h * vy1;
    hvmod = tsqrt(hu * hu + hv * hv);
  }
  return bingham_tau0_t(tauy0, K, hvmod, tmax_const(h, P.hfrict0), cand);
}

template <typename T>
__device__ __forceinline__ void ext_bingham(const DevParams &P, const DevDem &dem, const T &xx, const T &yy,
                                            const T &vx1, const T &vy1, const T &rho_i, const T &inx,
                                            const T &iny, const T &tauy0, const T &K, const T &tau_in, T &exo,
                                            T &eyo) {
  const double dt = P.dt;
  T h, hu, hv, hinv, vmod, hvmod, hvmodinv;
  h = rho_i;
  hu = h * vx1;
  hv = h * vy1;
  if (val(h) < P.comp) {
    h = tconst<T>(P.comp);
    hu = tconst<T>(0);
    hv = tconst<T>(0);
    hinv = tconst<T>(0);
    vmod = tconst<T>(0);
    hvmod = tconst<T>(0);
    hvmodinv = tconst<T>(0);
  } else {
    hinv = 1 / h;
    hvmod = tsqrt(hu * hu + hv * hv);
    vmod = hvmod * hinv;
    hvmodinv = (val(hvmod) >= 1.e-5) ? 1. / hvmod : tconst<T>(0);
  }
  T Zx, Zy;
  slope_at(dem, xx, yy, Zx, Zy);
  T exx = tconst<T>(0) + (-P.cgra) * Zx;
  T exy = tconst<T>(0) + (-P.cgra) * Zy;
  T Tau0 = tau_in, own;
  if (bingham_tau0_t(tauy0, K, hvmod, tmax_const(h, P.hfrict0), own)) Tau0 = own;
  const T cf4b = Tau0 / P.dens;
  T frictx = cf4b * hu * hvmodinv;
  T fricty = cf4b * hv * hvmodinv;
  const T TotFrict = tsqrt(frictx * frictx + fricty * fricty);
  T dcx = -frictx * hinv, dcy = -fricty * hinv;
  T ax = exx + inx, ay = exy + iny;
  if (val(vmod) > 1.e-5) {
    T vsx = vx1 + dt * (ax + dcx), vsy = vy1 + dt * (ay + dcy);
    T vdot = vsx * vx1 + vsy * vy1;
    if (val(vdot) < 0) {
      T tx = vx1 / vmod, ty = vy1 / vmod;
      T accelT = ax * tx + ay * ty;
      T deccelT = dcx * tx + dcy * ty;
      if (val(accelT) * val(deccelT) < 0) {
        dcx = -vx1 / dt - ax;
        dcy = -vy1 / dt - ay;
        frictx = -dcx * h;
        fricty = -dcy * h;
      }
    }
  } else {
    T accelT = tsqrt(ax * ax + ay * ay);
    if (val(accelT) > 1.e-5) {
      frictx = TotFrict * ax / accelT;
      fricty = TotFrict * ay / accelT;
      T fricT = tsqrt(frictx * frictx + fricty * fricty);
      if (val(accelT) <= val(fricT)) {
        frictx = ax;
        fricty = ay;
      }
    } else {
      frictx = tconst<T>(0);
      fricty = tconst<T>(0);
    }
    frictx = frictx * h;
    fricty = fricty * h;
  }
  exo = exx - frictx / h;
  eyo = exy - fricty / h;
}

// The Bingham stages of one particle given the Tau0 each stage resolved
// (tau[st], taken where the particle is not itself in the intermediate
// regime); cand[st] / ok[st]: the particle's own candidate at stage st.
#define BING_MAX_STAGES 4
template <typename T>
__device__ __forceinline__ void bing_stage_map(const DevParams &P, const DevDem &dem, const T &x0x, const T &x0y,
                                               const T &v0x, const T &v0y, const T &rf, const T &inx, const T &iny,
                                               const T &tauy0, const T &K, const T (&tau)[BING_MAX_STAGES],
                                               T &xox, T &xoy, T &vox, T &voy, T (&cand)[BING_MAX_STAGES],
                                               bool (&ok)[BING_MAX_STAGES], bool &mid) {
  const double dt = P.dt;
  T xx = x0x, xy = x0y, vx = v0x, vy = v0y;
  T ex = tconst<T>(0), ey = tconst<T>(0);
  T Dx = tconst<T>(0), Dy = tconst<T>(0), Dvx = tconst<T>(0), Dvy = tconst<T>(0);
  mid = false;
#pragma unroll
  for (int st = 0; st < BING_MAX_STAGES; st++) {
    ok[st] = false;
    cand[st] = tconst<T>(0);
  }
  for (int st = 0; st < P.nstages; st++) {
    const double f1 = P.f1[st];
    xx = x0x + f1 * dt * vx;
    xy = x0y + f1 * dt * vy;
    vx = v0x + f1 * dt * (ex + inx);
    vy = v0y + f1 * dt * (ey + iny);
    if (out_of_domain(dem, val(xx), val(xy))) {
      mid = true;
      break;
    }
    ok[st] = bingham_candidate_t(P, rf, vx, vy, tauy0, K, cand[st]);
    ext_bingham(P, dem, xx, xy, vx, vy, rf, inx, iny, tauy0, K, tau[st], ex, ey);
    const double f2 = P.f2[st];
    Dvx = Dvx + f2 * (ex + inx);
    Dvy = Dvy + f2 * (ey + iny);
    Dx = Dx + f2 * vx;
    Dy = Dy + f2 * vy;
  }
  xox = x0x + P.fin * Dx;
  xoy = x0y + P.fin * Dy;
  vox = v0x + P.fin * Dvx;
  voy = v0y + P.fin * Dvy;
}

// cubic-spline kernel with its gradient factor a (grad W = a d) and Hessian
// H = W'' dd^T/r^2 + (W'/r)(I - dd^T/r^2), from the squared separation dr
__device__ __forceinline__ void kernel_hess(double kc, double dr, double dx, double dy, double hsml, double &w,
                                            double &a, double &hxx, double &hxy, double &hyy) {
  double r, ir;
  root_nr(dr, r, ir);
  const double ih = rcp_nr(hsml), ih2 = ih * ih;
  const double q = r * ih, f = kc * ih2;
  double w2 = 0.0, a1 = 0.0;
  if (q <= 1.0) {
    w = f * (2. / 3. - q * q + 0.5 * (q * q * q));
    a = f * (-2. + 1.5 * q) * ih2;
    w2 = f * (-2. + 3. * q) * ih2;
  } else if (q <= 2.0) {
    const double t = 2. - q;
    w = f * (1.0 / 6.0) * (t * t * t);
    a = -(f * 0.5 * (t * t) * ih) * ir;
    w2 = f * t * ih2;
  } else {
    w = a = 0.0;
  }
  a1 = a;
  // unit vector; at r = 0 the Hessian is isotropic with W''(0) = W'/r(0)
  const double ux = r > 0 ? dx * ir : 0.0, uy = r > 0 ? dy * ir : 0.0;
  hxx = w2 * ux * ux + a1 * (1.0 - ux * ux);
  hxy = (w2 - a1) * ux * uy;
  hyy = w2 * uy * uy + a1 * (1.0 - uy * uy);
  if (!(r > 0)) {
    hxx = hyy = a1;
    hxy = 0.0;
  }
}

// per-step adjoint work arrays, in the step's sorted order
struct AdjWork {
  double2 *lxp, *lvpp, *lvp, *la, *lx, *lxn, *lvn;  // lambda of x', v'', v', a, x (pairs), x^n, v^n
  double *lS, *lm;
  int *act;  // takes part in XSPH
};

// recorded forward step (sorted order of that step)
struct TapeRec {
  double4 *nb;
  int *id, *rid, *sflag, *oflag;
  double *bt;  // Bingham: [2S][n] by caller id, each stage's resolved Tau0, then its candidates
  double *S;
  double2 *ain, *vp, *vn;  // vn: velocity at step start, caller order
  unsigned long long *mask;
  int *start;
  int start_cap;  // entries of start
  GridInfo *grid;
  int *coff;
};

__device__ __forceinline__ Cells tape_cells(const TapeRec &t) {
  Cells cl;
  cl.grid = t.grid;
  cl.coff = t.coff;
  cl.ntot = t.coff + 1;
  cl.start = t.start;
  return cl;
}

// K1: cotangents of the step outputs into sorted order
__global__ void k_adj_load(int n, int npoin0, TapeRec t, const double2 *LX, const double2 *LV, const double *LH,
                           AdjWork w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = t.id[i];
  w.lxp[i] = LX[k];
  w.lvpp[i] = LV[k];
  w.lS[i] = LH ? LH[k] : 0.0;
  w.lm[i] = 0.0;
  w.lx[i] = make_double2(0.0, 0.0);
  w.act[i] = (!t.sflag[i] && !t.oflag[i] && k >= npoin0) ? 1 : 0;
}

// Adjoint gathers run two lanes per particle: lane t sums the neighbours of
// ordinal parity t in their mask words (or stream positions t, t+2, ... of a
// full stream), and the lanes add their partial sums in a fixed order, so the
// result does not depend on the work-group size.  Loads of a batch are issued
// for every entry (an unused one reads the particle itself) and the sums take
// only the used entries.
#define ADJ_TM 2
// pieces per level when a checkpoint segment is longer than the records kept at once
#define ADJ_BRANCH 8

template <int NS>
__device__ __forceinline__ void adj_pair_sum(double (&a)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; k++) a[k] = a[k] + __shfl_xor(a[k], 1);
}

// K2: XSPH adjoint (gather over the step's neighbour masks)
__global__ __launch_bounds__(GF_BLOCK) void k_adj_xsph(int n, DevParams P, TapeRec t, AdjWork w) {
  int i, ln;
  team_index<ADJ_TM>(i, ln);
  if (i >= n) return;
  if (t.sflag[i]) {
    if (ln == 0) w.lvp[i] = w.lvpp[i];
    return;
  }
  const Cells cl = tape_cells(t);
  const int rr = t.rid[i];
  const GridInfo g = cl.grid[rr];
  const double4 qi = t.nb[i];
  const double2 pi_ = make_double2(qi.x, qi.y);
  const double hi = qi.z, mi = qi.w;
  const CandStream cs = cand_stream(g, cl.start + cl.coff[rr], pi_, hi, P.scale_k);
  const unsigned long long *mrow = t.mask + NB_WORDS * (long)i;
  const double2 vi = t.vp[i], li = w.lvpp[i];
  const double Si = t.S[i];
  const int acti = w.act[i];
  // csum, lSi, lmi, gxa, gya, lvx, lvy
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  auto term = [&](bool ok, const double4 &q, const double2 vj, const double2 lj, double Sj, int actj) {
    double dx, dy, dr, mh, wv, gx, gy;
    pair_geom(pi_, hi, q, dx, dy, dr, mh);
    sph_kernel_t<1>(P.kc, P.pi, dr, dx, dy, mh, wv, gx, gy);
    const double Ssum = Si + Sj;
    const double cij = (2 * q.w) * wv / Ssum, cji = (2 * mi) * wv / Ssum;
    const double dvx = vj.x - vi.x, dvy = vj.y - vi.y;
    const double pi_side = acti ? li.x * dvx + li.y * dvy : 0.0;       // lambda_i . (v_j - v_i)
    const double pj_side = actj ? -(lj.x * dvx + lj.y * dvy) : 0.0;    // lambda_j . (v_i - v_j)
    const double f = pi_side * 2.0 * q.w / Ssum + pj_side * 2.0 * mi / Ssum;
    const double c[7] = {acti ? cij : 0.0, -(pi_side * cij + pj_side * cji) / Ssum, pj_side * 2.0 * wv / Ssum,
                         f * gx, f * gy, actj ? cji * lj.x : 0.0, actj ? cji * lj.y : 0.0};
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = ok ? acc[k] + c[k] : acc[k];
  };
  if (stream_full(cs)) {
    stream_dense(cs, P.scale_k, i, ln, pi_, hi, t.nb, [&](int j, const double4 &q, double, double, double, double) {
      term(true, q, t.vp[j], w.lvpp[j], t.S[j], w.act[j]);
    });
  } else {
    unsigned long long mw[NB_WORDS];
    load_words(cs, mrow, mw, ~0ull);
#pragma unroll
    for (int k = 0; k < NB_WORDS; k++) mw[k] = ordinal_half(mw[k], ln);
    word_batches(cs, mw, i, [&](const int (&js)[NB_BATCH], const bool (&ok)[NB_BATCH]) {
      double4 q[NB_BATCH];
      double2 vj[NB_BATCH], lj[NB_BATCH];
      double Sj[NB_BATCH];
      int aj[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        q[u] = t.nb[js[u]];
        vj[u] = t.vp[js[u]];
        lj[u] = w.lvpp[js[u]];
        Sj[u] = t.S[js[u]];
        aj[u] = w.act[js[u]];
      }
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) term(ok[u], q[u], vj[u], lj[u], Sj[u], aj[u]);
    });
  }
  adj_pair_sum(acc);
  if (ln != 0) return;
  const double keep = acti ? 1.0 - 0.3 * acc[0] : 1.0;
  w.lvp[i] = make_double2(li.x * keep + 0.3 * acc[5], li.y * keep + 0.3 * acc[6]);
  w.lS[i] += 0.3 * acc[1];
  w.lm[i] += 0.3 * acc[2];
  const double2 lx = w.lx[i];
  w.lx[i] = make_double2(lx.x + 0.3 * acc[3], lx.y + 0.3 * acc[4]);
}

// K3: per-particle stage adjoint (Pr: the realisations' parameter blocks).
// The cotangents of the law's two parameters accumulate per particle, by
// caller id, over the steps (GP[id], GQ[id]); their sums per realisation are
// taken in id order at the end, so no order depends on the launch.
__global__ __launch_bounds__(GF_BLOCK) void k_adj_stage(int n, const DevParams *Pr, DevDem dem, TapeRec t,
                                                        AdjWork w, double *GP, double *GQ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 lx = w.lxp[i], lv = w.lvp[i];
  if (t.sflag[i]) {
    w.lxn[i] = lx;
    w.lvn[i] = lv;
    w.la[i] = make_double2(0.0, 0.0);
    return;
  }
  const DevParams &P = Pr[t.rid[i]];
  const int id = t.id[i];
  const double4 q = t.nb[i];
  const double2 v0 = t.vn[id], a = t.ain[i];
  const double Si = t.S[i];
  const double rf = Si < P.h_inf ? P.h_inf : Si;
  Dual X0 = dvar(q.x, 0), Y0 = dvar(q.y, 1), VX = dvar(v0.x, 2), VY = dvar(v0.y, 3), RF = dvar(rf, 4);
  Dual AX = dvar(a.x, 5), AY = dvar(a.y, 6), MU = dvar(P.tanfi8, 7), XI = dvar(P.xi_voellmy, 8);
  Dual xo, yo, vxo, vyo;
  bool mid;
  stage_map(P, dem, X0, Y0, VX, VY, RF, AX, AY, MU, XI, xo, yo, vxo, vyo, mid);
  double g[AD_N];
#pragma unroll
  for (int k = 0; k < AD_N; k++)
    g[k] = mid ? 0.0 : lx.x * xo.d[k] + lx.y * yo.d[k] + lv.x * vxo.d[k] + lv.y * vyo.d[k];
  w.lxn[i] = make_double2(g[0], g[1]);
  w.lvn[i] = make_double2(g[2], g[3]);
  if (Si >= P.h_inf) w.lS[i] += g[4];
  w.la[i] = make_double2(g[5], g[6]);
  GP[id] += g[7];
  GQ[id] += g[8];
}

// ---------------------------------------------------------------------------
// Bingham stage adjoint.  In stage st a particle outside the intermediate
// regime takes the Tau0 of the last particle before it (caller ids, its own
// realisation) whose candidate was valid, or the value carried from the
// previous scan.  With the resolved values tau_st held as inputs, a
// particle's stages are a map of its own inputs and of tau_0 .. tau_{S-1};
// the candidates it offers are outputs.  The cotangent of tau_st of a
// particle belongs to the candidate it was taken from: per stage, in reverse
// order, (1) k_adj_bing_tau gives every particle its lambda(tau_st) (its
// outputs' cotangents and those of its later candidates, forward-mode in the
// one direction tau_st), (2) a segmented reverse scan over ids adds each
// segment [valid k, next valid) to lambda(cand_st[k]); what precedes the
// realisation's first valid candidate belongs to the carried value, i.e.
// to the last valid candidate of the previous scan (or its carry), and is
// passed on in bcar[r]; (3) after the stages k_adj_stage_bing contracts the
// outputs' and the candidates' cotangents over the particle's own inputs.
// The tape holds tau_st and cand_st by id ([2S][n]: tau_0.., cand_0..).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool cand_valid(double c) { return c == c; }

__global__ __launch_bounds__(GF_BLOCK) void k_adj_bing_tau(int n, const DevParams *Pr, DevDem dem, TapeRec t,
                                                           AdjWork w, int sd, const double *lamCand,
                                                           double *lamTau) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int id = t.id[i];
  if (t.sflag[i]) {
    lamTau[id] = 0.0;
    return;
  }
  const DevParams &P = Pr[t.rid[i]];
  const int S = P.nstages;
  const double4 q = t.nb[i];
  const double2 v0 = t.vn[id], a = t.ain[i];
  const double Si = t.S[i];
  const double rf = Si < P.h_inf ? P.h_inf : Si;
  Dual tau[BING_MAX_STAGES];
#pragma unroll
  for (int st = 0; st < BING_MAX_STAGES; st++)
    tau[st] = st < S ? (st == sd ? dvar(t.bt[(long)st * n + id], 0) : dconst(t.bt[(long)st * n + id])) : dconst(0.0);
  Dual xo, yo, vxo, vyo, cand[BING_MAX_STAGES];
  bool ok[BING_MAX_STAGES], mid;
  bing_stage_map(P, dem, dconst(q.x), dconst(q.y), dconst(v0.x), dconst(v0.y), dconst(rf), dconst(a.x),
                 dconst(a.y), dconst(P.tauy0), dconst(P.constK), tau, xo, yo, vxo, vyo, cand, ok, mid);
  double g = 0.0;
  if (!mid) {
    const double2 lx = w.lxp[i], lv = w.lvp[i];
    g = lx.x * xo.d[0] + lx.y * yo.d[0] + lv.x * vxo.d[0] + lv.y * vyo.d[0];
#pragma unroll
    for (int st = 0; st < BING_MAX_STAGES; st++)
      if (st > sd && st < S && ok[st]) g += lamCand[(long)st * n + id] * cand[st].d[0];
  }
  lamTau[id] = g;
}

__global__ __launch_bounds__(GF_BLOCK) void k_adj_stage_bing(int n, const DevParams *Pr, DevDem dem, TapeRec t,
                                                             AdjWork w, const double *lamCand, double *GP,
                                                             double *GQ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 lx = w.lxp[i], lv = w.lvp[i];
  if (t.sflag[i]) {
    w.lxn[i] = lx;
    w.lvn[i] = lv;
    w.la[i] = make_double2(0.0, 0.0);
    return;
  }
  const DevParams &P = Pr[t.rid[i]];
  const int S = P.nstages;
  const int id = t.id[i];
  const double4 q = t.nb[i];
  const double2 v0 = t.vn[id], a = t.ain[i];
  const double Si = t.S[i];
  const double rf = Si < P.h_inf ? P.h_inf : Si;
  Dual tau[BING_MAX_STAGES];
#pragma unroll
  for (int st = 0; st < BING_MAX_STAGES; st++) tau[st] = dconst(st < S ? t.bt[(long)st * n + id] : 0.0);
  Dual xo, yo, vxo, vyo, cand[BING_MAX_STAGES];
  bool ok[BING_MAX_STAGES], mid;
  bing_stage_map(P, dem, dvar(q.x, 0), dvar(q.y, 1), dvar(v0.x, 2), dvar(v0.y, 3), dvar(rf, 4), dvar(a.x, 5),
                 dvar(a.y, 6), dvar(P.tauy0, 7), dvar(P.constK, 8), tau, xo, yo, vxo, vyo, cand, ok, mid);
  double g[AD_N];
#pragma unroll
  for (int k = 0; k < AD_N; k++) {
    double v = lx.x * xo.d[k] + lx.y * yo.d[k] + lv.x * vxo.d[k] + lv.y * vyo.d[k];
#pragma unroll
    for (int st = 0; st < BING_MAX_STAGES; st++)
      if (st < S && ok[st]) v += lamCand[(long)st * n + id] * cand[st].d[k];
    g[k] = mid ? 0.0 : v;
  }
  w.lxn[i] = make_double2(g[0], g[1]);
  w.lvn[i] = make_double2(g[2], g[3]);
  if (Si >= P.h_inf) w.lS[i] += g[4];
  w.la[i] = make_double2(g[5], g[6]);
  GP[id] += g[7];
  GQ[id] += g[8];
}

// Segmented reverse scan over the Bingham scan tiles (each tile a range of
// ids of one realisation; LV_ITEMS ids per thread).  A thread's "lead" is the
// sum of its ids before its first valid candidate (all of them if none);
// walking the threads (then the tiles) from the right, the carry into a range
// is what arrives from its right up to the next valid candidate.
__device__ void seg_thread_cin(const double *lam, const double *cand, int b, int e, double cin_tile,
                               double *cin_out, double *lead_out, bool *has_out) {
  __shared__ double lead[GF_BLOCK];
  __shared__ int has[GF_BLOCK];
  __shared__ double cin[GF_BLOCK];
  const int t = threadIdx.x;
  double l = 0.0;
  int hv = 0;
  for (int k = 0; k < LV_ITEMS; k++) {
    const int id = b + t * LV_ITEMS + k;
    if (id < e && !hv) {
      if (cand_valid(cand[id]))
        hv = 1;
      else
        l = l + lam[id];
    }
  }
  lead[t] = l;
  has[t] = hv;
  __syncthreads();
  if (t == 0) {
    double carry = cin_tile;
    int any = 0;
    for (int u = GF_BLOCK - 1; u >= 0; u--) {
      cin[u] = carry;
      carry = has[u] ? lead[u] : lead[u] + carry;
      any |= has[u];
    }
    if (lead_out) *lead_out = carry;
    if (has_out) *has_out = any != 0;
  }
  __syncthreads();
  *cin_out = cin[t];
}

__global__ __launch_bounds__(GF_BLOCK) void k_seg_reduce(const int4 *tiles, const double *lam, const double *cand,
                                                         double *tlead, int *thas) {
  const int4 tl = tiles[blockIdx.x];
  double cin, lead;
  bool has;
  seg_thread_cin(lam, cand, tl.y, tl.z, 0.0, &cin, &lead, &has);
  if (threadIdx.x == 0) {
    tlead[blockIdx.x] = lead;
    thas[blockIdx.x] = has ? 1 : 0;
  }
}

// one thread per realisation: the carry into each of its tiles, from its
// last tile (which receives the pending cotangent of the later scan's carry)
// to its first; what is left is the cotangent of this scan's carried value
__global__ void k_seg_top(int R, const int *rtile, const double *tlead, const int *thas, double *tcin,
                          double *bcar) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double carry = bcar[r];
  for (int b = rtile[r + 1] - 1; b >= rtile[r]; b--) {
    tcin[b] = carry;
    carry = thas[b] ? tlead[b] : tlead[b] + carry;
  }
  bcar[r] = carry;
}

__global__ __launch_bounds__(GF_BLOCK) void k_seg_down(const int4 *tiles, const double *lam, const double *cand,
                                                       const double *tcin, double *lamCand) {
  const int4 tl = tiles[blockIdx.x];
  double cin;
  seg_thread_cin(lam, cand, tl.y, tl.z, tcin[blockIdx.x], &cin, nullptr, nullptr);
  const int b0 = tl.y + threadIdx.x * LV_ITEMS;
  double acc = cin;
  for (int k = LV_ITEMS - 1; k >= 0; k--) {
    const int id = b0 + k;
    if (id >= tl.z) continue;
    acc = lam[id] + acc;
    if (cand_valid(cand[id])) {
      lamCand[id] = acc;
      acc = 0.0;
    } else {
      lamCand[id] = 0.0;
    }
  }
}

// K4: pressure-force and density adjoints (gathers), then the cotangents of
// x^n, v^n and the mass back to caller order
__global__ __launch_bounds__(GF_BLOCK) void k_adj_pairs(int n, DevParams P, TapeRec t, AdjWork w, double2 *LX,
                                                        double2 *LV, double *LM) {
  int i, ln;
  team_index<ADJ_TM>(i, ln);
  if (i >= n) return;
  const int k = t.id[i];
  if (t.sflag[i]) {
    if (ln == 0) {
      LX[k] = w.lxn[i];
      LV[k] = w.lvn[i];
    }
    return;
  }
  const Cells cl = tape_cells(t);
  const int rr = t.rid[i];
  const GridInfo g = cl.grid[rr];
  const double4 qi = t.nb[i];
  const double2 pi_ = make_double2(qi.x, qi.y);
  const double hi = qi.z, mi = qi.w;
  const CandStream cs = cand_stream(g, cl.start + cl.coff[rr], pi_, hi, P.scale_k);
  const unsigned long long *mrow = t.mask + NB_WORDS * (long)i;
  const double2 lai = w.la[i];
  const double lSi = w.lS[i];
  const double grav = 2.0 * P.cgra05;
  double acc[3] = {0.0, 0.0, 0.0};  // lxx, lxy, lmi
  auto term = [&](bool ok, const double4 &q, const double2 laj, double lSj) {
    double dx, dy, dr, mh, wv, a, hxx, hxy, hyy;
    pair_geom(pi_, hi, q, dx, dy, dr, mh);
    kernel_hess(P.kc, dr, dx, dy, mh, wv, a, hxx, hxy, hyy);
    // pressure force: -g sum_j H_ij (m_j la_i - m_i la_j); mass: sum_j la_j . (g grad W_ij)
    const double bx = q.w * lai.x - mi * laj.x, by = q.w * lai.y - mi * laj.y;
    double cx = -grav * (hxx * bx + hxy * by);
    double cy = -grav * (hxy * bx + hyy * by);
    double cm = grav * a * (laj.x * dx + laj.y * dy);
    // density: sum_j (lS_i m_j + lS_j m_i) grad W_ij; mass: sum_j lS_j W_ij
    const double c = (lSi * q.w + lSj * mi) * a;
    acc[0] = ok ? (acc[0] + cx) + c * dx : acc[0];
    acc[1] = ok ? (acc[1] + cy) + c * dy : acc[1];
    acc[2] = ok ? (acc[2] + cm) + lSj * wv : acc[2];
  };
  if (stream_full(cs)) {
    stream_dense(cs, P.scale_k, i, ln, pi_, hi, t.nb, [&](int j, const double4 &q, double, double, double, double) {
      term(true, q, w.la[j], w.lS[j]);
    });
  } else {
    unsigned long long mw[NB_WORDS];
    load_words(cs, mrow, mw, ~0ull);
#pragma unroll
    for (int kk = 0; kk < NB_WORDS; kk++) mw[kk] = ordinal_half(mw[kk], ln);
    word_batches(cs, mw, i, [&](const int (&js)[NB_BATCH], const bool (&ok)[NB_BATCH]) {
      double4 q[NB_BATCH];
      double2 laj[NB_BATCH];
      double lSj[NB_BATCH];
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) {
        q[u] = t.nb[js[u]];
        laj[u] = w.la[js[u]];
        lSj[u] = w.lS[js[u]];
      }
#pragma unroll
      for (int u = 0; u < NB_BATCH; u++) term(ok[u], q[u], laj[u], lSj[u]);
    });
  }
  adj_pair_sum(acc);
  if (ln != 0) return;
  const double lmi = acc[2] + lSi * sph_self_w(1, P.pi, hi);
  const double2 lxn = w.lxn[i], lxp = w.lx[i];
  LX[k] = make_double2(lxn.x + lxp.x + acc[0], lxn.y + lxp.y + acc[1]);
  LV[k] = w.lvn[i];
  LM[k] += w.lm[i] + lmi;
}

// Deposit misfit L = sum_nodes (R - T)^2 of the IDW depth raster R
// (OutputRes_WIR) and its cotangents with respect to the particles' x and h
__global__ void k_misfit_nodes(long np, const double *acc5, const double *target, double *res, double *partial) {
  __shared__ double red[GF_BLOCK / 64];
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (g < np) {
    const double ws = acc5[5 * g + 4];
    const double R = ws >= 1.e-6 ? acc5[5 * g] / ws : acc5[5 * g];
    const double e = R - target[g];
    res[2 * g] = R;
    res[2 * g + 1] = e;
    l = e * e;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < GF_BLOCK / 64; k++) s += red[k];
    partial[blockIdx.x] = s;
  }
}

__global__ void k_misfit_vjp(int n, DevDem dem, CurView s, const double *acc5, const double *res, double2 *LX,
                             double *LH, int rsel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (rsel >= 0 && s.rid[i] != rsel)) return;
  const int k = s.id[i];
  if (s.flag[i]) {
    LX[k] = make_double2(0.0, 0.0);
    LH[k] = 0.0;
    return;
  }
  const double2 p = s.pos[i];
  const double h = s.h[i];
  const double distg = sqrt(dem.dx * dem.dx + dem.dy * dem.dy);
  int ipx = (int)((p.x - dem.xmin) / dem.dx + 1), ipy = (int)((p.y - dem.ymin) / dem.dy + 1);
  int inx0 = ipx - 1 > 1 ? ipx - 1 : 1, iny0 = ipy - 1 > 1 ? ipy - 1 : 1;
  int inx1 = ipx + 2 < dem.nx ? ipx + 2 : dem.nx, iny1 = ipy + 2 < dem.ny ? ipy + 2 : dem.ny;
  double lh = 0.0, lx = 0.0, ly = 0.0;
  for (int inx = inx0; inx <= inx1; inx++)
    for (int iny = iny0; iny <= iny1; iny++) {
      long g = (long)(iny - 1) * dem.nx + inx - 1;
      const double ws = acc5[5 * g + 4];
      if (!(ws >= 1.e-6)) continue;
      const double R = res[2 * g], e2 = 2.0 * res[2 * g + 1];
      double gx = dem.xmin + (inx - 1) * dem.dx, gy = dem.ymin + (iny - 1) * dem.dy;
      double a = p.x - gx, b = p.y - gy;
      const double rho = sqrt(a * a + b * b);
      const double dist = fmax(1.e-6, rho) / distg;
      lh += e2 / (dist * ws);
      if (rho > 1.e-6) {
        // d(1/dist)/dx = -distg (x - node) / rho^3
        const double c = e2 * (h - R) / ws * (-distg / (rho * rho * rho));
        lx += c * a;
        ly += c * b;
      }
    }
  LX[k] = make_double2(lx, ly);
  LH[k] = lh;
}

// ---------------------------------------------------------------------------
// initialisation helpers
// ---------------------------------------------------------------------------
__global__ void k_domain_flag(int n, DevDem dem, const double2 *pos, int *flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 p = pos[i];
  if (out_of_domain(dem, p.x, p.y)) flag[i] = 1;
}

// Input_SW: h = summed depth (6:1440), then pwp initialisation (6:1705-1852)
__global__ void k_init_pwp(int n, DevParams P, double alfs, double pwp0rel, int resum, const double *S,
                           SortedSet s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (resum && !s.flag[i]) s.h[i] = S[i];
  double hh = s.h[i] * alfs;
  if (P.icpwp == 1) {
    s.u[i] = pwp0rel * hh;
  } else if (P.icpwp == 3) {
    s.uaux[i] = pwp0rel * hh;
    s.u[i] = s.uaux[i];
  }
}

// ---------------------------------------------------------------------------
// Pore pressure on a column of N = npwp nodes per particle (icpwp = 2):
// Get_Pwp_SW for npwp > 1 (6:10966-11174), single phase, no boundary zones.
// Each particle's column, bed (node 0) to surface, is consolidated
// explicitly (FTCS) over the saturated depth h alphaS in ntpwp sub-steps of
// dt_sph, with the loading dh/ntpwp (1 - z/h), clamps that depend on the
// height, and the bed / surface conditions (1 value, 2 gradient).  The
// columns are node-major ([N][n]) so that the lanes' accesses to one node
// are contiguous; the sweep runs in place with the previous node's old value
// kept in a register.  The sub-step count is capped at PWP_MAX_SUBSTEPS so
// that every lane ends (the cap is reached only for a vanishing depth).
// ---------------------------------------------------------------------------
#define PWP_MAX_SUBSTEPS (1 << 16)

__global__ void k_pwp_column(int n, const DevParams *Pr, DevDem dem, SortedSet s, NewSet o, const int *stopr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = s.rid[i];
  const DevParams &P = Pr[r];
  const int N = P.npwp;
  const double *src = s.col + i;
  double *col = o.col + i;
  if (stopr && stopr[r]) {  // a stopped case does not change
    for (int k = 0; k < N; k++) col[(long)k * n] = src[(long)k * n];
    return;
  }
  if (o.flag[i]) {  // Uaux(:, ipoin) = 0 out of the domain (6:10977-10980)
    for (int k = 0; k < N; k++) col[(long)k * n] = 0.0;
    o.uaux[i] = 0.0;
    o.u[i] = 0.0;
    return;
  }
  for (int k = 0; k < N; k++) col[(long)k * n] = src[(long)k * n];
  const double Cv = P.Cv, alphaS = P.alphaS, dt = P.dt;
  const double hh = o.h[i], hpwp = hh * alphaS, dhh = o.dh[i];
  const double dxpwp = hpwp / (N - 1), dxpwp2 = dxpwp * dxpwp;
  double dtpwp = (Cv > 1.e-20) ? dxpwp2 / (2 * Cv) : dxpwp2 / (2.e-20);
  const double q = ceil(dt / dtpwp);
  const int ntpwp = (q >= 1.0 && q <= PWP_MAX_SUBSTEPS) ? (int)q : (q >= 1.0 ? PWP_MAX_SUBSTEPS : 1);
  dtpwp = dt / (double)ntpwp;
  const double dhhstep = dhh / ntpwp, dpwp_dn_step = 0.0 / ntpwp;
  const double r_ = Cv * dtpwp / dxpwp2;
  const double fmax_ = P.fpwp_max, fmin_ = P.fpwp_min;
  // the bed condition of a pore-pressure zone of the terrain (basal type 3:
  // LBc1 = int(basal_props(7)), VBc1 = basal_props(8), 6:10987-11001), at
  // the step's final position
  int bt1 = P.bc_t1;
  double bv1 = P.bc_v1;
  if (dem.zkinds & 4) {
    const double2 pp = o.pos[i];
    const int z = which_zone(dem, P, pp.x, pp.y, 3);
    if (z >= 0) {
      bt1 = (int)dem.zones[z].p[6];
      bv1 = dem.zones[z].p[7];
    }
  }
  for (int it = 0; it < ntpwp; it++) {
    double prev = col[0], cur = col[(long)n];
    for (int k = 1; k < N - 1; k++) {
      const double next = col[(long)(k + 1) * n];
      const double zz = dxpwp * k, zzhh = hh - zz, xi = zz / hh;
      const double sumdhh = dhhstep * (1. - xi);
      double v = cur + r_ * (prev - 2. * cur + next) + sumdhh + dpwp_dn_step;
      if (v > (zzhh * fmax_)) v = (zzhh * fmax_);
      else if (v < (zzhh * fmin_)) v = (zzhh * fmin_);
      col[(long)k * n] = v;
      prev = cur;
      cur = next;
    }
    double b = col[0];
    if (bt1 == 1) b = bv1 * hh * alphaS;
    else if (bt1 == 2) b = (4. * col[(long)n] - col[(long)2 * n] - 2. * bv1 * dxpwp) / 3.;
    if (b > hh) b = hh;
    col[0] = b;  // before the surface condition, which reads node 0 when N = 3
    double t = col[(long)(N - 1) * n];
    if (P.bc_t2 == 1) t = P.bc_v2 * hh;
    else if (P.bc_t2 == 2) t = (4. * col[(long)(N - 2) * n] - col[(long)(N - 3) * n] + 2. * P.bc_v2 * dxpwp) / 3.;
    if (t > 0.0) t = 0.0;
    col[(long)(N - 1) * n] = t;
  }
  o.uaux[i] = col[0];
  o.u[i] = col[0];  // u = Uaux(1, :) (7:956-958)
}

// Initialize_Pwp for icpwp = 2 (6:1769-1803): the initial column profile of
// every particle from its (re-summed) depth
__global__ void k_init_column(int n, const DevParams *Pr, int pwp0_type, double pwp0rel, double xirel, SortedSet s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int N = Pr[s.rid[i]].npwp;
  const double hh = s.h[i], pi = 4. * atan(1.0);
  for (int k = 0; k < N; k++) {
    double xi, v = 0.0;
    if (pwp0_type == 0) {
      v = pwp0rel * hh;
    } else if (pwp0_type == 1) {
      xi = 1. * k / (N - 1);
      v = (1 - xi) * pwp0rel * hh;
    } else if (pwp0_type == 2) {
      xi = (k + 1 - 1.) / (N - 1.);
      v = cos(pi * xi / 2) * pwp0rel * hh;
    } else if (pwp0_type == 3) {
      v = (k == 0) ? pwp0rel * hh : 0.0;
    } else {
      xi = (k + 1 - 1.) / (N - 1.);
      v = (xi < xirel) ? hh * pwp0rel * (1 - xi) : 0.0;
    }
    s.col[(long)k * n + i] = v;
  }
  s.uaux[i] = s.col[i];
  s.u[i] = s.col[i];
}

// Input_SW's case variable 401 (6:1453-1479): every particle's pore
// pressure u and Uaux(:) times its realisation's factor (the state the next
// step sorts from, nw, and the sorted set it is in the order of)
__global__ void k_scale_pwp(int n, const double *f, SortedSet s, NewSet o, int ncol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = f[s.rid[i]];
  o.u[i] = o.u[i] * c;
  o.uaux[i] = o.uaux[i] * c;
  s.u[i] = s.u[i] * c;
  s.uaux[i] = s.uaux[i] * c;
  for (int k = 0; k < ncol; k++) {
    o.col[(long)k * n + i] = o.col[(long)k * n + i] * c;
    s.col[(long)k * n + i] = s.col[(long)k * n + i] * c;
  }
}

// new-set view of a sorted set (used to re-sort the initial state)
__global__ void k_copy_to_new(int n, SortedSet s, NewSet o, int ncol) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  o.pos[i] = s.pos[i];
  o.h[i] = s.h[i];
  o.hsml[i] = s.hsml[i];
  o.u[i] = s.u[i];
  o.uaux[i] = s.uaux[i];
  o.flag[i] = s.flag[i];
  for (int k = 0; k < ncol; k++) o.col[(long)k * n + i] = s.col[(long)k * n + i];
}

// FD nodal derivatives (Fdiff_derivs, 2:2812-2877)
__global__ void k_fdiff(int nx, int ny, double dxg, double dyg, const double *Z, const int *in_model, double2 *zxy,
                        double *top) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long np = (long)nx * ny;
  if (k >= np) return;
  int ix = (int)(k % nx), iy = (int)(k / nx);
  double zx = 0, zy = 0, t[6] = {0, 0, 0, 0, 0, 0};
  bool inner = ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2;
  if (inner && (in_model == nullptr || in_model[k])) {
    double del2x = 2. * dxg, del2y = 2. * dyg, delx2 = dxg * dxg, dely2 = dyg * dyg, delxy = dxg * dyg;
    double zA = Z[k], zL = Z[k - 1], zR = Z[k + 1], zT = Z[k + nx], zB = Z[k - nx];
    double zTL = Z[k + nx - 1], zTR = Z[k + nx + 1], zBL = Z[k - nx - 1], zBR = Z[k - nx + 1];
    zx = (zR - zL) / del2x;
    zy = (zT - zB) / del2y;
    double rt = sqrt(1. + zx * zx + zy * zy);
    double zxx = (zL - 2. * zA + zR) / delx2;
    double zyy = (zB - 2. * zA + zT) / dely2;
    double zxy2 = (zTR - zBR - zTL + zBL) / (4. * delxy);
    t[0] = 1. + zx * zx;
    t[1] = zx * zy;
    t[2] = 1. + zy * zy;
    t[3] = zxx / rt;
    t[4] = zxy2 / rt;
    t[5] = zyy / rt;
  }
  zxy[k] = make_double2(zx, zy);
  if (top)
    for (int c = 0; c < 6; c++) top[6 * k + c] = t[c];
}

__global__ void k_topol_split(long np, const double *topol9, double2 *zxy, double *top) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  zxy[k] = make_double2(topol9[9 * k + 1], topol9[9 * k + 2]);
  if (top)
    for (int c = 0; c < 6; c++) top[6 * k + c] = topol9[9 * k + 3 + c];
}

// un-permute to caller order
__global__ void k_unsort(int n, CurView s, double2 *x, double2 *v, double *h, double *u, double *uaux,
                         double *hsml, int *flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int d = s.id[i];
  x[d] = s.pos[i];
  v[d] = s.vel[i];
  h[d] = s.h[i];
  u[d] = s.u[i];
  uaux[d] = s.uaux[i];
  hsml[d] = s.hsml[i];
  flag[d] = s.flag[i];
}

// pore-pressure columns between the caller's layout [n][N] (by id) and the
// current state's node-major slots
__global__ void k_col_get(int n, int N, const int *id, const double *col, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long d = (long)id[i] * N;
  for (int k = 0; k < N; k++) out[d + k] = col[(long)k * n + i];
}

__global__ void k_col_set(int n, int N, const int *id, const double *in, double *col_new, double *col_sorted,
                          double *uaux, double *u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long d = (long)id[i] * N;
  for (int k = 0; k < N; k++) {
    col_new[(long)k * n + i] = in[d + k];
    col_sorted[(long)k * n + i] = in[d + k];
  }
  uaux[i] = in[d];
  u[i] = in[d];
}

__global__ void k_unsort_mass(int n, CurView s, double *mass) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mass[s.id[i]] = s.mass[i];
}

__global__ void k_nonfinite(int n, CurView s, int *bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 p = s.pos[i], v = s.vel[i];
  double q = p.x + p.y + v.x + v.y + s.h[i] + s.u[i];
  if (!isfinite(q)) atomicOr(bad, 1);
}

__global__ void k_count_active(int n, const int *flag, int *count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned a = (i < n && flag[i] == 0) ? 1u : 0u;
  a = wave_sum_u32(a);
  if ((threadIdx.x & 63) == 0 && a) atomicAdd(count, (int)a);
}

// ---------------------------------------------------------------------------
// a15 adaptive dt_sph: Adaptive_dt_sph_SW (6:6521-6664), called by
// Driver_Flow before each time_integration_sph (8:1590).  Modes
// (ic_adapt_dt_sph): 1 hsml/(3c), 2 hsml/(3(c+|v|)), 4 |dx_g|/|v|,
// 10 min of 1 and 4, with c = sqrt(g h) and h clamped to h_inf in place.
// The minimum is over the active particles of each realisation; positive
// doubles order as their bit patterns, so the minimum is an integer
// atomicMin and its value does not depend on the reduction order.
// ---------------------------------------------------------------------------
#define DT_INIT 0x4202a05f20000000ull  // bit pattern of 1.e10, the reference's dtmin start

__global__ __launch_bounds__(GF_BLOCK) void k_dt_min(int n, int mode, double h_inf, double cgra, double dxg,
                                                     const double2 *vel, double *h, const double *hsml,
                                                     const int *flag, const int *rid, const int *stopr,
                                                     unsigned long long *dtm) {
  __shared__ unsigned long long wmin[GF_BLOCK / 64];
  const int b0 = blockIdx.x * blockDim.x;
  const int i = b0 + threadIdx.x;
  unsigned long long key = ~0ull;
  if (i < n && !flag[i] && !(stopr && stopr[rid[i]])) {
    const double2 v = vel[i];
    const double vv = sqrt(v.x * v.x + v.y * v.y);
    double dm = 1.e10;
    if (mode != 4) {
      double hh = h[i];
      if (hh < h_inf) {
        hh = h_inf;
        h[i] = hh;  // the reference keeps the clamped depth (6:6554)
      }
      double cc = fmax(1.e-6, sqrt(cgra * hh));
      if (mode == 2) cc = cc + vv;
      dm = fmin(hsml[i] / (3 * cc), dm);
    }
    if (mode == 4 || mode == 10) dm = fmin(dxg / vv, dm);
    key = (unsigned long long)__double_as_longlong(dm);
  }
  // slots are grouped by realisation: a block whose first and last slot
  // share one reduces in LDS and issues one atomic
  const int rf = rid[b0], rl = rid[min(n, b0 + (int)blockDim.x) - 1];
  if (rf != rl) {
    if (key != ~0ull) atomicMin(&dtm[rid[i]], key);
    return;
  }
  key = wave_min_u64(key);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = wmin[w] < m ? wmin[w] : m;
    if (m != ~0ull) atomicMin(&dtm[rf], m);
  }
}

// ic_adapt_dt_sph = 7, the one-dimensional free surface (6:6597-6607): the
// particles in the caller's order, each with the next one of its
// realisation, give 0.5 |x_(k+1) - x_k| / v where v = |vx_(k+1) + vx_k| / 2
// exceeds 1e-6 -- every particle, inside the domain or not, no depth clamp.
// k_dt_fs_scatter puts (x, vx) and the realisation in caller order first.
__global__ void k_dt_fs_scatter(int n, const int *id, const double2 *pos, const double2 *vel, const int *rid,
                                double2 *xv, int *rk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = id[i];
  xv[k] = make_double2(pos[i].x, vel[i].x);
  rk[k] = rid[i];
}

__global__ void k_dt_fs(int n, const double2 *xv, const int *rk, const int *stopr, unsigned long long *dtm) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k + 1 >= n) return;
  const int r = rk[k];
  if (rk[k + 1] != r || (stopr && stopr[r])) return;
  const double2 a = xv[k], b = xv[k + 1];
  const double Le = fabs(b.x - a.x);
  const double vv = fabs(b.y + a.y) / 2.;
  if (vv > 1.e-6) {
    const double cc = 0.5 * Le / vv;
    atomicMin(&dtm[r], (unsigned long long)__double_as_longlong(cc));
  }
}

// New dt_sph of every realisation (6:6655-6660), its RK weight, time_sph
// (8:1617; with case control k_case_pre adds the step's dt instead), and the
// minimum slots reset for the next step.  Stopped realisations keep theirs.
__global__ void k_dt_apply(int R, int nstages, DevParams *Pr, unsigned long long *dtm, double *tm, const int *stopr,
                           int add_time) {
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    const double dtmin = __longlong_as_double((long long)dtm[r]);
    dtm[r] = DT_INIT;
    if (stopr && stopr[r]) continue;
    const double dt = Pr[r].dt;
    const double nd = dtmin < dt ? dtmin : fmin(1.1 * dt, dtmin);
    Pr[r].dt = nd;
    Pr[r].fin = nstages == 4 ? nd / 6. : nd / 2.;
    if (add_time) tm[r] = tm[r] + nd;
  }
}

// ---------------------------------------------------------------------------
// Case control: the step loop of Driver_Flow_23 (8:1444-1548) that the
// reference runs for every case of a case table or Monte Carlo set, here for
// every realisation of a batch at once, on the device.  Per step:
//   k_case_pre   dt_sph cut so that the step ends on the next plot time
//                (8:1481-1488), t_plot_reset, time_sph and time advanced
//   (the step; a stopped realisation is copied unchanged)
//   k_case_due   plot time reached (8:1498): the realisation's pixels are
//                accumulated (Get_MC_pixel_Vars_SW) by k_mc_pixel
//   k_case_out, k_case_sum   out-of-domain counts and sum of |vx| of the
//                active particles, per realisation
//   k_case_stop  dt_sph restored after a plot (8:1508), Stop_case_SW
//                (6:20114-20177) and time >= time_end (8:1528)
// ---------------------------------------------------------------------------
struct CaseState {
  double tplot;    // t_plot_reset
  double tglob;    // time
  double dt_keep;  // dt_sph before the cut to the plot time
  double vsum;     // sum of |vx| over the active particles after the step
  int nsteps;      // itimestep: loop iterations done
  int pad[3];
};

struct CaseCtl {
  double vstop, timestop, dtimestop, dt, time_end, time_plot;
};

__global__ void k_case_pre(int R, int nstages, DevParams *Pr, double *tm, CaseState *cs, const int *stopr,
                           double time_plot) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || stopr[r]) return;
  double dt = Pr[r].dt;
  cs[r].dt_keep = dt;
  if (time_plot > 0.0) {
    const double tpr = cs[r].tplot + dt;
    if (tpr >= time_plot) {
      const int nt = (int)(tpr / time_plot);
      dt = (double)nt * time_plot - cs[r].tplot;
    }
    cs[r].tplot = cs[r].tplot + dt;
  }
  Pr[r].dt = dt;
  Pr[r].fin = nstages == 4 ? dt / 6. : dt / 2.;
  tm[r] = tm[r] + dt;
  cs[r].tglob = cs[r].tglob + dt;
}

__global__ void k_case_due(int R, const CaseState *cs, const int *stopr, double time_plot, int *due, int *outc) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  due[r] = (!stopr[r] && time_plot > 0.0 && cs[r].tplot >= time_plot) ? 1 : 0;
  outc[r] = 0;
}

// out-of-domain particles per realisation (integer counts: order-free)
__global__ void k_case_out(int n, const int *flag, const int *rid, int *outc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) atomicAdd(&outc[rid[i]], 1);
}

// one block per realisation: sum of |vx| over its active particles (its
// cells' slot range), in a fixed order
__global__ __launch_bounds__(GF_BLOCK) void k_case_sum(const int *coff, const int *start, const double2 *vel,
                                                       const int *flag, const int *stopr, CaseState *cs) {
  __shared__ double part[GF_BLOCK];
  const int r = blockIdx.x;
  if (stopr[r]) return;
  const int b = start[coff[r]], e = start[coff[r + 1]];
  double acc = 0.0;
  for (int i = b + (int)threadIdx.x; i < e; i += GF_BLOCK)
    if (!flag[i]) acc = acc + fabs(vel[i].x);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int o = GF_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) cs[r].vsum = part[0];
}

__global__ void k_case_stop(int R, int nstages, DevParams *Pr, const double *tm, CaseState *cs, int *stopr,
                            const int *due, const int *outc, const int *cnt_r, CaseCtl ctl) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || stopr[r]) return;
  const int it = cs[r].nsteps + 1;
  cs[r].nsteps = it;
  if (due[r]) {
    const double dk = cs[r].dt_keep;
    Pr[r].dt = dk;
    Pr[r].fin = nstages == 4 ? dk / 6. : dk / 2.;
    cs[r].tplot = 0.0;
  }
  bool stop = false;
  if (it > 1) {
    const double v = cs[r].vsum / (double)cnt_r[r];
    if (v < ctl.vstop && it > 10) stop = true;
    if (outc[r] > cnt_r[r] / 5) stop = true;
    if (tm[r] > ctl.timestop) stop = true;
    if (ctl.dt < ctl.dtimestop) stop = true;
  }
  if (cs[r].tglob >= ctl.time_end) stop = true;
  if (stop) stopr[r] = 1;
}

// pair enumeration (i<j by caller index) for the neighbour-set export
__global__ void k_pairs(int n, int skf, int scale_k, Cells cl, SortedSet s, long long cap,
                        unsigned long long *count, int *ij) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || s.flag[i]) return;
  const int rr_ = s.rid[i];
  const GridInfo g = cl.grid[rr_];
  const int *start = cl.start + cl.coff[rr_];
  const double2 pi_ = s.pos[i];
  const double hi = s.hsml[i];
  const int idi = s.id[i];
  for_candidates(g, start, pi_, [&](int j) {
    if (j == i) return;
    int idj = s.id[j];
    if (idj <= idi) return;
    double2 pj = s.pos[j];
    double dx, dy, r, mh;
    if (pair_pred(scale_k, pi_, pj, hi, s.hsml[j], dx, dy, r, mh)) {
      unsigned long long k = atomicAdd(count, 1ull);
      if ((long long)k < cap) {
        ij[2 * k] = idi;
        ij[2 * k + 1] = idj;
      }
    }
  });
}

// a14 (i): IDW particle -> node scatter (6:14312-14440), soil particles
// rsel >= 0: the particles of realisation rsel only
__global__ void k_raster_idw(int n, DevDem dem, CurView s, double *acc5, int rsel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || s.flag[i] || (rsel >= 0 && s.rid[i] != rsel)) return;
  double2 p = s.pos[i];
  double2 v = s.vel[i];
  double h = s.h[i], uu = s.u[i];
  double distg = sqrt(dem.dx * dem.dx + dem.dy * dem.dy);
  int ipx = (int)((p.x - dem.xmin) / dem.dx + 1), ipy = (int)((p.y - dem.ymin) / dem.dy + 1);
  int inx0 = ipx - 1 > 1 ? ipx - 1 : 1, iny0 = ipy - 1 > 1 ? ipy - 1 : 1;
  int inx1 = ipx + 2 < dem.nx ? ipx + 2 : dem.nx, iny1 = ipy + 2 < dem.ny ? ipy + 2 : dem.ny;
  for (int inx = inx0; inx <= inx1; inx++)
    for (int iny = iny0; iny <= iny1; iny++) {
      long gidx = (long)(iny - 1) * dem.nx + inx - 1;
      double gx = dem.xmin + (inx - 1) * dem.dx, gy = dem.ymin + (iny - 1) * dem.dy;
      double a = p.x - gx, b = p.y - gy;
      a = a * a;
      b = b * b;
      double dist = fmax(1.e-6, sqrt(a + b));
      dist = dist / distg;
      unsafeAtomicAdd(&acc5[5 * gidx + 0], h / dist);
      unsafeAtomicAdd(&acc5[5 * gidx + 1], v.x / dist);
      unsafeAtomicAdd(&acc5[5 * gidx + 2], v.y / dist);
      unsafeAtomicAdd(&acc5[5 * gidx + 3], (uu / h) / dist);
      unsafeAtomicAdd(&acc5[5 * gidx + 4], 1.00 / dist);
    }
}

__global__ void k_raster_norm(long np, const double *acc5, double *out, double *wsum) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  double ws = acc5[5 * g + 4];
  wsum[g] = ws;
  for (int k = 0; k < 4; k++) out[4 * g + k] = ws >= 1.e-6 ? acc5[5 * g + k] / ws : acc5[5 * g + k];
}

// a14 (ii): MC pixel accumulation (6:20717-20833)
__device__ __forceinline__ void atomic_max_pos(double *addr, double v) {
  // values are >= 0: the IEEE bit pattern orders like the unsigned integer
  atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}

// stride > 0: per-realisation accumulators, realisation r at offset r*stride nodes
__global__ void k_mc_pixel(int n, DevDem dem, CurView s, long stride, double *acc4, int *cnt, const int *due) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)__lane_id();
  // lanes with a cell take part; the others carry keys of their own
  long key = -1 - lane, g1 = 0;
  double hh = 0.0, vv = 0.0;
  if (i < n && !s.flag[i] && !(due && !due[s.rid[i]])) {
    const long base = stride * s.rid[i];
    const double2 p = s.pos[i];
    const double2 v = s.vel[i];
    int ipx = (int)((p.x - dem.xmin) / dem.dx + 1), ipy = (int)((p.y - dem.ymin) / dem.dy + 1);
    if (ipx == dem.nx && fabs(dem.xmax - p.x) <= 1.e-3) ipx--;
    if (ipy == dem.ny && fabs(dem.ymax - p.y) <= 1.e-3) ipy--;
    if (!(ipx < 1 || ipy < 1 || ipx > dem.nx - 1 || ipy > dem.ny - 1)) {
      g1 = base + (long)(ipy - 1) * dem.nx + ipx - 1;
      key = g1;
      hh = s.h[i];
      vv = sqrt(v.x * v.x + v.y * v.y);  // **0.5: flang emits a square root
    }
  }
  // particles of a wave that share a cell (they arrive in cell order) are
  // combined first: one set of atomics per run of equal cells
  const WaveRun w = wave_run(key, lane);
  const int end = lane + w.len;
  double sh = hh, sv = vv, mh = hh, mv = vv;
  int c = 1;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double th = __shfl_down(sh, o), tv = __shfl_down(sv, o), xh = __shfl_down(mh, o), xv = __shfl_down(mv, o);
    const int tc = __shfl_down(c, o);
    if (lane + o < end) {
      sh += th;
      sv += tv;
      mh = xh > mh ? xh : mh;
      mv = xv > mv ? xv : mv;
      c += tc;
    }
  }
  if (lane != w.head || key < 0) return;
  const long nodes[4] = {g1, g1 + 1, g1 + 1 + dem.nx, g1 + dem.nx};
  for (int k = 0; k < 4; k++) {
    const long g = nodes[k];
    atomicAdd(&cnt[g], c);
    unsafeAtomicAdd(&acc4[4 * g + 0], sh);  // hardware FP64 add
    unsafeAtomicAdd(&acc4[4 * g + 2], sv);
    atomic_max_pos(&acc4[4 * g + 1], mh);
    atomic_max_pos(&acc4[4 * g + 3], mv);
  }
}

// Monte Carlo statistics per node over the realisations (Norm_MC_pixel_Vars_SW,
// 6:20838-20908): each realisation that reached the node contributes its
// per-node averages (sums over plot times / visits) and maxima; the mean and
// standard deviation (variance below 1e-6 taken as 0) are over those
// realisations.  Also the maxima over all realisations and the fraction of
// realisations whose maximum depth exceeds hthr.
__global__ void k_mc_hazard(long np, int R, const double *acc4, const int *cnt, double hthr, double *stats,
                            int *nprob, double *maxs, double *pexc) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0}, mh = 0.0, mv = 0.0;
  int np_r = 0, nexc = 0;
  for (int r = 0; r < R; r++) {
    const long k = (long)r * np + g;
    const int c = cnt[k];
    if (c == 0) continue;
    np_r++;
    const double4 a = reinterpret_cast<const double4 *>(acc4)[k];
    const double o[4] = {a.x / c, a.y, a.z / c, a.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      s1[q] += o[q];
      s2[q] += o[q] * o[q];
    }
    mh = fmax(mh, a.y);
    mv = fmax(mv, a.w);
    nexc += a.y > hthr ? 1 : 0;
  }
  const int nd = np_r > 0 ? np_r : 1;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double mean = s1[q] / nd, sig2 = s2[q] / nd;
    double t = sig2 - mean * mean;
    if (t < 1.e-6) t = 0.0;
    stats[8 * g + 2 * q] = mean;
    stats[8 * g + 2 * q + 1] = sqrt(t);
  }
  nprob[g] = np_r;
  maxs[2 * g] = mh;
  maxs[2 * g + 1] = mv;
  pexc[g] = (double)nexc / R;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct gfsph_ctx {
  int n = 0;
  DevParams P{};
  DevDem dem{};
  int check_finite = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // device buffers
  SortedSet set[2]{};
  NewSet nw{};
  double *S = nullptr, *Pp = nullptr;
  double4 *nbp = nullptr;  // per slot (x, y, pressure, mass) from the density pass, read by the force pass
  unsigned long long *mask = nullptr;
  int4 *srec = nullptr;  // the step's candidate-stream records [n][2]
  double2 *ain = nullptr;  // pressure force per particle
  int *key = nullptr, *rank = nullptr, *perm = nullptr, *cnt = nullptr, *start = nullptr, *bsum = nullptr;
  int cap_cells = 0, scan_blocks = 0;
  GridInfo *grid = nullptr;
  BBoxAcc *acc = nullptr, *part = nullptr;
  BBoxAcc *boxpart = nullptr;  // batches: per-realisation box partials [R][box_chunks]
  int box_chunks = 1;
  int gridN = 0;  // blocks of the per-particle kernels (multiple of 8)
  hipGraphExec_t graph[2] = {nullptr, nullptr};   // one step from set[k]
  hipGraphExec_t graphG[2] = {nullptr, nullptr};  // GRAPH_STEPS steps from set[k]
  // work-group size of each gather kernel (density, force, XSPH), chosen by
  // gfsph_autotune; all sizes give identical results
  int gbs[3] = {GATHER_BLOCK, GATHER_BLOCK, GATHER_BLOCK};
  // lanes per particle of each gather (1 or 2; the same bits either way):
  // 2 below TEAM1_MIN_N particles, else 1; gfsph_autotune may change them
  int gteam[3] = {2, 2, 2};
  int *bad = nullptr;
  double2 *zxy = nullptr;
  double *top = nullptr;
  int has_zone1 = 0;  // a terrain zone of geometry 1 reads the current Z
  double *raster5 = nullptr, *mc_acc = nullptr;
  int *mc_cnt = nullptr;
  BinghamState bs{};
  int lv_blocks = 0;
  int cur = 0;
  // realisations (1 for a single run): per-realisation grids, cell offsets,
  // slot ranges and parameter blocks
  int R = 1;
  int *coff = nullptr;
  int2 *rng = nullptr;
  DevParams *Pr = nullptr;
  bool first_build = true;
  // per-realisation Monte Carlo pixel accumulators [R][npoig] (lazily allocated)
  double *mcr_acc = nullptr;
  int *mcr_cnt = nullptr;
  // adjoint: checkpoints of the whole device state every ckpt steps, the
  // records of one recomputed segment, cotangents in caller order
  struct Snap {
    SortedSet s;
    NewSet o;
    int cur;
    double *carry;  // Bingham: each realisation's carried Tau0
    int2 *rng;      // batches: each realisation's slot range in the saved sorted set
  };
  int ad_ckpt = 0, ad_nsteps = 0;
  int ad_tape = 0;         // most steps recorded at once (0: ckpt); longer segments get a second level
  int ad_cells = 0;        // largest cell count the forward pass's checkpoints saw
  int *ad_over = nullptr;  // device: cells a recorded step needed beyond its buffer (0: none)
  std::vector<std::vector<Snap>> snapsL;  // checkpoints inside a segment, per level
  std::vector<Snap> snaps;  // [nseg + 1]: segment starts and the final state
  std::vector<TapeRec> tape;
  AdjWork aw{};
  double2 *LX = nullptr, *LV = nullptr;
  double *LM = nullptr, *LH = nullptr;
  double *GP = nullptr, *GQ = nullptr;  // parameter cotangents per particle, by caller id
  // Bingham adjoint: lambda of each particle's resolved Tau0 and candidates,
  // the segmented-scan tile sums, the pending cotangent of each carried Tau0;
  // tape_bt: where the step being recorded writes its stages' Tau0
  double *lamTau = nullptr, *lamCand = nullptr, *tlead = nullptr, *tcin = nullptr, *bcar = nullptr;
  int *thas = nullptr;
  double *tape_bt = nullptr;
  std::vector<DevParams> Ph;  // the realisations' parameter blocks (host copy)
  double *mass0 = nullptr;    // nfric = 3: masses before any erosion, by caller id
  double2 *fs_xv = nullptr;   // ic_adapt_dt_sph = 7: (x, vx) by caller id
  int *fs_rk = nullptr;       //   and the realisation of each caller id
  // time step: ic_adapt_dt_sph (0: fixed dt_sph); with adaptation dt_sph and
  // time_sph of every realisation live on the device (Pr[r].dt, tm, dtm the
  // per-step minima), otherwise time_sph is summed here from dt_h
  int adapt = 0;
  double dxg = 0.0;
  double *tm = nullptr;
  unsigned long long *dtm = nullptr;
  std::vector<double> time_h, dt_h;
  // case control (Driver_Flow_23 per realisation): thresholds, per-realisation
  // loop state, stop flags, plot flags, out-of-domain counts, sizes
  int cases = 0, case_pixels = 0;
  std::vector<int> counts_h;  // particles of each realisation
  CaseCtl cctl{};
  CaseState *cs = nullptr;
  int *stopr = nullptr, *due = nullptr, *outc = nullptr, *cnt_r = nullptr;
  // erosion (some realisation has ic_erosion): the terrain's erosion state
  int eros = 0;
  // pressure re-evaluated in every stage (some realisation has ic_K0 or ic_cosTheta)
  int staged = 0;
  // the basal-law class every realisation shares (LAW_FRIC5: nfric = 5 without
  // cosTheta; else LAW_ANY), selecting the specialised RK kernel
  int law = LAW_ANY;
  ErosGrid eg{};
  std::vector<void *> allocs;
};

static Cells cells_of(const gfsph_ctx *c) {
  Cells cl;
  cl.grid = c->grid;
  cl.coff = c->coff;
  cl.ntot = c->coff + c->R;
  cl.start = c->start;
  return cl;
}

static CurView cur_view(const gfsph_ctx *c) {
  const SortedSet &s = c->set[c->cur];
  CurView v;
  v.pos = c->nw.pos;
  v.vel = s.vel;
  v.h = c->nw.h;
  v.mass = s.mass;
  v.hsml = c->nw.hsml;
  v.u = c->nw.u;
  v.uaux = c->nw.uaux;
  v.flag = c->nw.flag;
  v.id = s.id;
  v.rid = s.rid;
  return v;
}

static int fail(gfsph_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(ctx, GFSPH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

template <typename T>
static int dalloc(gfsph_ctx *c, T **p, size_t count) {
  void *q = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(c, GFSPH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = (T *)q;
  return GFSPH_OK;
}

static inline int nblk(long n) { return (int)((n + GF_BLOCK - 1) / GF_BLOCK); }

static int derive_params(const gfsph_params *p, DevParams &P, std::string &err) {
  const double *c = p->cst;
  memset(&P, 0, sizeof(P));
  P.nfric = (int)(c[5] + 0.01);
  P.if_frictional = 0;
  P.if_V2byR = 0;
  if ((P.nfric >= 5 && P.nfric <= 10) || P.nfric == 26 || P.nfric == 27) {
    P.if_frictional = 1;
    if (P.nfric == 7 || P.nfric == 8 || P.nfric == 10 || P.nfric == 26 || P.nfric == 27) P.if_V2byR = 1;
  }
  // 201-203 read the porosity n_DF of the two-phase debris-flow model
  // (6:9491-9536, 6:9728-9742), which single-phase runs do not allocate
  const int ok_nfric[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 22, 23, 24, 25, 26, 27};
  bool ok = false;
  for (int v : ok_nfric) ok = ok || (v == P.nfric);
  if (!ok) {
    err = "basal law nfric=" + std::to_string(P.nfric) + " is outside the implemented set";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (P.nfric == 23) {
    // the exponents the reference accepts (6:9578-9601; others stop it)
    const double m = c[8];
    if (m == 1) P.hb_m = 1;
    else if (m == 2) P.hb_m = 2;
    else if (fabs(m - 0.333333) <= 1.e-3 || m == 0.5) P.hb_m = 3;
    else {
      err = "nfric = 23 takes the exponent const(8) = 1, 2, 1/2 or 1/3 only";
      return GFSPH_ERR_ARG;
    }
    const double Cm = m / ((m + 1) * (2 * m + 1));
    P.hb_A = (1. / Cm) * pow(c[7] / c[6], 1. / m);
  }
  P.costheta = (int)(c[16] + 0.01);
  if (P.costheta != 0 && P.costheta != 1) {
    err = "ic_cosTheta (const(16)) must be 0 or 1";
    return GFSPH_ERR_UNSUPPORTED;
  }
  // the reference reads ic_K0 for frictional laws only (6:1373-1390)
  P.ic_k0 = (p->ic_k0 == 1 && P.if_frictional) ? 1 : 0;
  P.k0a = p->k0_active;
  P.k0p = p->k0_pasive;
  if ((P.costheta || P.ic_k0) && P.nfric == 4) {
    err = "ic_cosTheta = 1 with the Bingham law (nfric = 4) is outside the implemented path";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (c[20] != 0.0 || c[34] != 0.0) {
    err = "only ic_fct_dens = 0 and ic_hrelSat = 0 are implemented";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (c[36] > 0.01 || c[43] > 0) {
    err = "crushing and end-solid laws are outside the implemented path";
    return GFSPH_ERR_UNSUPPORTED;
  }
  // erosion: Input_SW maps const(4) > 1e-4 to law 1 with K = const(4)
  // (6:1180-1184); ic_erosion = const(21) > 1e-9 (6:1288-1290)
  {
    double c21 = c[21], c22 = c[22];
    if (c[4] > 1.e-4) {
      c21 = 1;
      c22 = c[4];
    }
    P.eros_on = c21 > 1.e-9;
    P.eros_law = (int)(c21 + 0.1);
    P.keros = c22;
    P.eros_c23 = c[23];
    P.eros_c24 = c[24];
    P.eros_c25 = c[25];
    P.denss = c[17];
    P.densw_e = c[18];
    P.dens_bar_eff = c[2] - (c[18] < 1.e-3 ? 1000.0 : c[18]);  // 6:9001-9008
    P.slip_fact = 1.0;                                          // nconst = 110 > 45 (6:8672-8679)
    if (c[47] > 1.e-6) P.slip_fact = 2 * c[47] / (1. + c[47]);
    P.avail0 = (c[48] > 1.e-6) ? c[48] : 999.;                 // topol(15,:) (6:1294-1301)
    if (P.eros_on) {
      if ((P.eros_law == 2 || P.eros_law == 3) && (c[17] <= 1.e-3 || c[18] <= 1.e-3)) {
        err = "erosion laws 2 and 3 need const(17) and const(18) (the reference stops here)";
        return GFSPH_ERR_UNSUPPORTED;
      }
      if (P.eros_law == 2 && !P.if_frictional) {
        err = "erosion law 2 reads the frictional law's tan(phi): it needs a frictional nfric";
        return GFSPH_ERR_UNSUPPORTED;
      }
      if (P.nfric == 4) {
        err = "erosion with the Bingham law (nfric = 4) is outside the implemented path";
        return GFSPH_ERR_UNSUPPORTED;
      }
    }
  }
  if (p->summation_density != 1 || p->nor_density != 0) {
    err = "summation density without normalisation is required";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (!(p->pa_sph == 1 || p->pa_sph == 2 || p->pa_sph == 4)) {
    err = "pa_sph must be 1, 2 or 4";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (p->skf < 1 || p->skf > 4) {
    err = "skf must be 1..4";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (p->sle < 0 || p->sle > 3) {
    err = "sle must be 0, 1, 2 or 3";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (!(p->icpwp == 0 || p->icpwp == 1 || p->icpwp == 2 || p->icpwp == 3)) {
    err = "icpwp must be 0, 1, 2 or 3";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (!(p->integrator == 4 || p->integrator == 2)) {
    err = "integrator must be 4 (RK4) or 2 (RK2)";
    return GFSPH_ERR_UNSUPPORTED;
  }
  if (!(p->dt > 0) || c[15] <= 1e-8 || c[10] <= 1e-8) {
    err = "dt, comp (const(15)) and hfrict0 (const(10)) must be positive";
    return GFSPH_ERR_ARG;
  }
  if (p->icpwp == 3 && !(c[13] > 0)) {
    err = "icpwp = 3 needs a positive dissipation rate Bfact (const(13))";
    return GFSPH_ERR_ARG;
  }
  if ((P.nfric == 4 || P.nfric == 23 || P.nfric == 24 || P.nfric == 25) && c[6] < 1.e-32) {
    err = "the yield stress const(6) must exceed 1e-32 (6:9547-9550)";
    return GFSPH_ERR_ARG;
  }
  if (p->icpwp == 2) {
    if (p->nuaux < 3 || p->nuaux > 1024) {
      err = "icpwp = 2 needs a column of 3 to 1024 nodes (nuaux)";
      return GFSPH_ERR_ARG;
    }
    for (int k = 0; k < 2; k++)
      if (!(p->pwp_bc_type[k] == 1 || p->pwp_bc_type[k] == 2)) {
        err = "pore-pressure column conditions must be of type 1 (value) or 2 (gradient)";
        return GFSPH_ERR_UNSUPPORTED;
      }
    if (p->pwp_bc_val[0] < -900) {
      err = "the drained-bed value (BC_pwp_val(1) < -900) is outside the implemented path";
      return GFSPH_ERR_UNSUPPORTED;
    }
    if (p->pwp0_type < 0 || p->pwp0_type > 4) {
      err = "pwp0_type must be 0..4";
      return GFSPH_ERR_ARG;
    }
  }
  P.icpwp = p->icpwp;
  P.nuaux = (p->icpwp == 3 || p->icpwp == 2) ? 1 : 0;  // a bed value Uaux(1,:) drives the friction
  P.npwp = (p->icpwp == 2) ? p->nuaux : 0;
  P.bc_t1 = p->pwp_bc_type[0];
  P.bc_t2 = p->pwp_bc_type[1];
  P.bc_v1 = p->pwp_bc_val[0];
  P.bc_v2 = p->pwp_bc_val[1];
  P.pa_sph = p->pa_sph;
  P.skf = p->skf;
  P.scale_k = (p->skf == 1 || p->skf == 4) ? 2 : 3;
  P.sle = p->sle;
  P.av_vel = p->average_velocity;
  P.npoin0 = p->npoin0;
  P.dt = p->dt;
  P.h_inf = p->h_inf;
  P.cgra = c[1];
  P.cgra05 = c[1] * 0.5;
  P.dens = c[2];
  P.xi_voellmy = c[3];
  P.constMu_CF = (c[3] < 0) ? -(25. * c[3] / 4.) : 0.0;
  P.K_Bagnold = c[3];
  P.tauy0 = c[6];
  P.constK = c[7];
  P.visco = c[8];
  P.tanfi8 = c[9];
  P.hfrict0 = c[10];
  P.Bfact = c[13];
  P.comp = c[15];
  P.alphaS = c[14];
  P.hrelpw2 = (double)((float)c[14] * (float)c[14]);
  P.c73 = 7. / 3.;
  P.fact_dens = 1.0;
  P.fact_pwp_max = 1.0;
  P.fact_pwp_min = 0.0;
  P.fpwp_max = 1.0;
  P.fpwp_min = 0.0;
  P.c3 = c[3];
  P.c6 = c[6];
  P.c7 = c[7];
  P.c8 = c[8];
  P.c47 = c[47];
  P.pi = 4.0 * atan(1.0);
  P.Cv = 4 * c[13] / (P.pi * P.pi);  // Get_Pwp_SW (6:10926)
  P.kc = p->skf == 1 ? (15.0 / 7.0) / P.pi : (7.0 / 4.0) / P.pi;
  if (p->integrator == 4) {
    const double f1[4] = {0., 0.5, 0.5, 1.0}, f2[4] = {1., 2., 2., 1.0};
    P.nstages = 4;
    for (int k = 0; k < 4; k++) {
      P.f1[k] = f1[k];
      P.f2[k] = f2[k];
    }
    P.fin = p->dt / 6.;
  } else {
    P.nstages = 2;
    P.f1[0] = 0.;
    P.f1[1] = 1.;
    P.f2[0] = 1.;
    P.f2[1] = 1.;
    P.fin = p->dt / 2.;
  }
  return GFSPH_OK;
}

static int alloc_sets(gfsph_ctx *c, int n) {
  int rc;
  for (int k = 0; k < 2; k++) {
    SortedSet &s = c->set[k];
    if ((rc = dalloc(c, &s.nb, n)) || (rc = dalloc(c, &s.pos, n)) || (rc = dalloc(c, &s.vel, n)) || (rc = dalloc(c, &s.h, n)) ||
        (rc = dalloc(c, &s.mass, n)) || (rc = dalloc(c, &s.hsml, n)) || (rc = dalloc(c, &s.u, n)) ||
        (rc = dalloc(c, &s.uaux, n)) || (rc = dalloc(c, &s.flag, n)) || (rc = dalloc(c, &s.id, n)) ||
        (rc = dalloc(c, &s.rid, n)))
      return rc;
  }
  NewSet &o = c->nw;
  if ((rc = dalloc(c, &o.pos, n)) || (rc = dalloc(c, &o.vel, n)) || (rc = dalloc(c, &o.h, n)) ||
      (rc = dalloc(c, &o.hsml, n)) || (rc = dalloc(c, &o.u, n)) || (rc = dalloc(c, &o.uaux, n)) ||
      (rc = dalloc(c, &o.flag, n)))
    return rc;
  if (c->P.npwp) {  // pore-pressure columns, node-major, cleared
    const size_t nc = (size_t)c->P.npwp * n;
    if ((rc = dalloc(c, &c->set[0].col, nc)) || (rc = dalloc(c, &c->set[1].col, nc)) || (rc = dalloc(c, &o.col, nc)) ||
        (rc = dalloc(c, &o.dh, n)))
      return rc;
    for (double *q : {c->set[0].col, c->set[1].col, o.col})
      HIPCHK(c, hipMemsetAsync(q, 0, sizeof(double) * nc, c->stream));
  }
  if ((rc = dalloc(c, &c->mask, NB_WORDS * (size_t)n)) || (rc = dalloc(c, &c->ain, n))) return rc;
  if ((rc = dalloc(c, &c->srec, 2 * (size_t)n))) return rc;
  if ((rc = dalloc(c, &c->S, n)) || (rc = dalloc(c, &c->Pp, n)) || (rc = dalloc(c, &c->nbp, n)) || (rc = dalloc(c, &c->key, n)) ||
      (rc = dalloc(c, &c->rank, n)) || (rc = dalloc(c, &c->perm, n)))
    return rc;
  // a single run caps its grid at CAP_PER_PART n + 4096 cells, realisation r of a batch
  // at CAP_PER_PART n_r + 64
  c->cap_cells = CAP_PER_PART * n + 64 * c->R + 4096;
  c->scan_blocks = (c->cap_cells + 2 + SCAN_TILE - 1) / SCAN_TILE;
  if ((rc = dalloc(c, &c->cnt, (size_t)c->cap_cells + 2)) || (rc = dalloc(c, &c->start, (size_t)c->cap_cells + 2)) ||
      (rc = dalloc(c, &c->bsum, (size_t)c->scan_blocks)) || (rc = dalloc(c, &c->grid, c->R)) ||
      (rc = dalloc(c, &c->coff, c->R + 1)) || (rc = dalloc(c, &c->rng, c->R)) || (rc = dalloc(c, &c->Pr, c->R)) ||
      (rc = dalloc(c, &c->acc, 1)) || (rc = dalloc(c, &c->bad, 1)))
    return rc;
  c->gridN = ((nblk(n) + 7) / 8) * 8;
  if ((rc = dalloc(c, &c->part, c->gridN))) return rc;
  if (c->R > 1) {  // a realisation's slot range never exceeds its particle count
    int mx = 1;
    for (int v : c->counts_h) mx = v > mx ? v : mx;
    c->box_chunks = (mx + BOX_CHUNK - 1) / BOX_CHUNK;
    if ((rc = dalloc(c, &c->boxpart, (size_t)c->R * c->box_chunks))) return rc;
  }
  // the cell counts start zero; every build clears the ones it used
  HIPCHK(c, hipMemsetAsync(c->cnt, 0, sizeof(int) * ((size_t)c->cap_cells + 2), c->stream));
  return GFSPH_OK;
}

static int alloc_bingham(gfsph_ctx *c, int n) {
  int rc;
  BinghamState &b = c->bs;
  // scan tiles of LV_TILE ids inside each realisation (ids are stacked in
  // realisation order)
  std::vector<int4> tiles;
  std::vector<int> rtile(c->R + 1, 0);
  int first = 0;
  for (int r = 0; r < c->R; r++) {
    const int nr = c->R == 1 ? n : c->counts_h[r];
    rtile[r] = (int)tiles.size();
    for (int b0 = first; b0 < first + nr; b0 += LV_TILE) tiles.push_back(make_int4(r, b0, std::min(b0 + LV_TILE, first + nr), 0));
    first += nr;
  }
  rtile[c->R] = (int)tiles.size();
  b.ntiles = (int)tiles.size();
  c->lv_blocks = b.ntiles;
  if ((rc = dalloc(c, &b.in, n)) || (rc = dalloc(c, &b.x, n)) || (rc = dalloc(c, &b.v, n)) ||
      (rc = dalloc(c, &b.ex, n)) || (rc = dalloc(c, &b.u, n)) || (rc = dalloc(c, &b.du, n)) ||
      (rc = dalloc(c, &b.D, 5 * (size_t)n)) || (rc = dalloc(c, &b.fl, n)) || (rc = dalloc(c, &b.cand, n)) ||
      (rc = dalloc(c, &b.tau, n)) || (rc = dalloc(c, &b.carry, c->R)) || (rc = dalloc(c, &b.bsum, b.ntiles)) ||
      (rc = dalloc(c, &b.tiles, b.ntiles)) || (rc = dalloc(c, &b.rtile, c->R + 1)))
    return rc;
  std::vector<double> zero(c->R, 0.0);  // Tau0 before any intermediate-regime particle
  HIPCHK(c, hipMemcpy(b.carry, zero.data(), sizeof(double) * c->R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.tiles, tiles.data(), sizeof(int4) * b.ntiles, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.rtile, rtile.data(), sizeof(int) * (c->R + 1), hipMemcpyHostToDevice));
  return GFSPH_OK;
}

#define SKF_DISPATCH(KERNEL, ...)                   \
  do {                                              \
    switch (c->P.skf) {                             \
      case 1: KERNEL<1><<<__VA_ARGS__; break;       \
      case 2: KERNEL<2><<<__VA_ARGS__; break;       \
      case 3: KERNEL<3><<<__VA_ARGS__; break;       \
      default: KERNEL<4><<<__VA_ARGS__; break;      \
    }                                               \
  } while (0)

// gathers of team size TMV (1 or 2 lanes per particle) for each kernel smoothing function
#define SKF_TM_DISPATCH(KERNEL, TMV, ...)                \
  do {                                                   \
    switch (c->P.skf) {                                  \
      case 1: KERNEL<1, TMV><<<__VA_ARGS__; break;       \
      case 2: KERNEL<2, TMV><<<__VA_ARGS__; break;       \
      case 3: KERNEL<3, TMV><<<__VA_ARGS__; break;       \
      default: KERNEL<4, TMV><<<__VA_ARGS__; break;      \
    }                                                    \
  } while (0)

#define PFORCE_PA(SKFV, PAV)                                                                                  \
  do {                                                                                                        \
    if (tm == 1)                                                                                              \
      k_pforce_m<SKFV, PAV, 1><<<xcd_grid((long)c->n, bsz), bsz, 0, c->stream>>>(PFORCE_ARGS);                \
    else                                                                                                      \
      k_pforce_m<SKFV, PAV, 2><<<xcd_grid((long)c->n * 2, bsz), bsz, 0, c->stream>>>(PFORCE_ARGS);            \
  } while (0)
#define PFORCE_CASE(SKFV)                      \
  switch (c->P.pa_sph) {                       \
    case 1: PFORCE_PA(SKFV, 1); break;         \
    case 2: PFORCE_PA(SKFV, 2); break;         \
    default: PFORCE_PA(SKFV, 4); break;        \
  }

// blocks of bsz threads covering n particles, rounded up to a multiple of 8 (xcd_block)
static inline int xcd_grid(long n, int bsz) { return (int)((((n + bsz - 1) / bsz) + 7) / 8 * 8); }

static void launch_pforce(gfsph_ctx *c, SortedSet &s) {
  const int bsz = c->gbs[1], tm = c->gteam[1];
#define PFORCE_ARGS c->n, c->P, cells_of(c), c->mask, s, c->S, c->Pp, c->ain, c->srec, c->staged ? nullptr : c->nbp
  switch (c->P.skf) {
    case 1: PFORCE_CASE(1) break;
    case 2: PFORCE_CASE(2) break;
    case 3: PFORCE_CASE(3) break;
    default: PFORCE_CASE(4) break;
  }
#undef PFORCE_ARGS
}

// the pore-pressure columns after the stages (icpwp = 2)
static void launch_column(gfsph_ctx *c, SortedSet &s) {
  if (c->P.npwp)
    k_pwp_column<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->Pr, c->dem, s, c->nw,
                                                          c->cases ? c->stopr : nullptr);
}

static void launch_pforce(gfsph_ctx *c, SortedSet &s);

// RK stages of one step on the sorted set s (pressure force in c->ain).
static void launch_stages(gfsph_ctx *c, SortedSet &s) {
  const int n = c->n;
  hipStream_t st = c->stream;
  const int nb = nblk(n);
  if (c->staged) {
    BinghamState &b = c->bs;
    const int *stp = c->cases ? c->stopr : nullptr;
    k_stg_init<<<nb, GF_BLOCK, 0, st>>>(n, s, b);
    for (int k = 0; k < c->P.nstages; k++) {
      k_stg_update<<<nb, GF_BLOCK, 0, st>>>(n, k, c->Pr, c->dem, s, c->S, c->ain, b, c->Pp, stp);
      if (c->P.ic_k0)
        SKF_DISPATCH(k_stg_k0, nb, GF_BLOCK, 0, st>>>(n, c->Pr, cells_of(c), c->mask, s, c->S, b, c->Pp, stp));
      launch_pforce(c, s);
      if (c->eros)
        k_stg_ext<true><<<nb, GF_BLOCK, 0, st>>>(n, k, c->Pr, c->dem, s, c->S, c->ain, b, stp);
      else
        k_stg_ext<false><<<nb, GF_BLOCK, 0, st>>>(n, k, c->Pr, c->dem, s, c->S, c->ain, b, stp);
    }
    BBoxAcc *part = c->R == 1 ? c->part : nullptr;
    if (c->eros) {
      k_stg_finish<true><<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, c->ain, b, c->nw, part, stp,
                                                        c->eg);
      const long tot = c->eg.np * c->R;
      k_erode_nodes<<<(int)((tot + GF_BLOCK - 1) / GF_BLOCK), GF_BLOCK, 0, st>>>(tot, c->eg);
    } else {
      k_stg_finish<false><<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, c->ain, b, c->nw, part, stp,
                                                         c->eg);
    }
    launch_column(c, s);
    return;
  }
  if (c->P.nfric != 4) {
    if (c->eros) {
      k_rk<true, LAW_ANY><<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, c->ain, c->nw,
                                                c->R == 1 ? c->part : nullptr, c->cases ? c->stopr : nullptr, c->eg);
      const long tot = c->eg.np * c->R;
      k_erode_nodes<<<(int)((tot + GF_BLOCK - 1) / GF_BLOCK), GF_BLOCK, 0, st>>>(tot, c->eg);
    } else if (c->law == LAW_FRIC5) {
      k_rk<false, LAW_FRIC5><<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, c->ain, c->nw,
                                                            c->R == 1 ? c->part : nullptr,
                                                            c->cases ? c->stopr : nullptr, c->eg);
    } else {
      k_rk<false, LAW_ANY><<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, c->ain, c->nw,
                                                          c->R == 1 ? c->part : nullptr, c->cases ? c->stopr : nullptr,
                                                          c->eg);
    }
    launch_column(c, s);
    return;
  }
  BinghamState &b = c->bs;
  k_bingham_init<<<nb, GF_BLOCK, 0, st>>>(n, s, c->ain, b);
  for (int k = 0; k < c->P.nstages; k++) {
    k_bingham_stage<<<nb, GF_BLOCK, 0, st>>>(n, 0, k, c->Pr, c->dem, s, c->S, b);
    k_lv_reduce<<<b.ntiles, GF_BLOCK, 0, st>>>(b.tiles, b.cand, b.bsum);
    k_lv_top<<<(c->R + 63) / 64, 64, 0, st>>>(c->R, b.rtile, b.bsum, b.carry);
    k_lv_down<<<b.ntiles, GF_BLOCK, 0, st>>>(b.tiles, b.cand, b.bsum, b.tau);
    if (c->tape_bt) {  // a step recorded for the adjoint (launched directly, not in a graph)
      const int S = c->P.nstages;
      (void)hipMemcpyAsync(c->tape_bt + (size_t)k * n, b.tau, sizeof(double) * n, hipMemcpyDeviceToDevice, st);
      (void)hipMemcpyAsync(c->tape_bt + (size_t)(S + k) * n, b.cand, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           st);
    }
    k_bingham_stage<<<nb, GF_BLOCK, 0, st>>>(n, 1, k, c->Pr, c->dem, s, c->S, b);
  }
  k_bingham_finish<<<c->gridN, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, s, c->S, b, c->nw,
                                                  c->R == 1 ? c->part : nullptr);
  launch_column(c, s);
}

// Neighbour build into set[dst] from (nw, set[src]).  A single run takes its
// grid from the bounding-box partials of the step that wrote nw; a batch
// computes every realisation's box over its slot range (from the previous
// sort, or from the caller's layout at the first build).
static int neighbour_build(gfsph_ctx *c, int src, int dst) {
  const int n = c->n, R = c->R;
  hipStream_t st = c->stream;
  const int nb = nblk(n);
  const int *ntot = c->coff + R;
  if (R == 1) {
    k_grid<<<1, GF_BLOCK, 0, st>>>(c->gridN, c->part, c->acc, c->grid, c->coff, c->P.scale_k, c->cap_cells);
  } else {
    if (!c->first_build) k_rng<<<(R + GF_BLOCK - 1) / GF_BLOCK, GF_BLOCK, 0, st>>>(R, c->coff, c->start, c->rng);
    k_bbox_chunks<<<dim3(R, c->box_chunks), GF_BLOCK, 0, st>>>(c->rng, c->nw.pos, c->nw.hsml, c->nw.flag,
                                                                c->boxpart);
    k_grid_real<<<(R + GF_BLOCK - 1) / GF_BLOCK, GF_BLOCK, 0, st>>>(R, c->box_chunks, c->rng, c->boxpart,
                                                                    c->P.scale_k, c->grid);
    k_coff<<<1, GF_BLOCK, 0, st>>>(R, c->grid, c->coff);
  }
  c->first_build = false;
  k_keys<<<nb, GF_BLOCK, 0, st>>>(n, c->nw.pos, c->nw.flag, c->set[src].rid, c->grid, c->coff, ntot, c->key,
                                  c->rank, c->cnt);
  k_scan_reduce<<<c->scan_blocks, GF_BLOCK, 0, st>>>(ntot, c->cnt, c->bsum);
  k_scan_down<<<c->scan_blocks, GF_BLOCK, 0, st>>>(ntot, c->cnt, c->bsum, c->start, n);
  k_scatter<<<nb, GF_BLOCK, 0, st>>>(n, c->key, c->rank, c->start, c->perm);
  k_reorder<<<nb, GF_BLOCK, 0, st>>>(n, ntot, c->perm, c->key, c->start, c->nw, c->set[src], c->set[dst],
                                     c->P.npwp);
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

__global__ void k_copy_vel(int n, const double2 *src, double2 *dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static void launch_density(gfsph_ctx *c, SortedSet &s) {
  const int bsz = c->gbs[0];
  if (c->gteam[0] == 1)
    SKF_TM_DISPATCH(k_density_m, 1, xcd_grid((long)c->n, bsz), bsz, 0,
                    c->stream>>>(c->n, c->P, cells_of(c), s, c->S, c->Pp, c->mask, c->srec, c->nbp));
  else
    SKF_TM_DISPATCH(k_density_m, 2, xcd_grid((long)c->n * 2, bsz), bsz, 0,
                    c->stream>>>(c->n, c->P, cells_of(c), s, c->S, c->Pp, c->mask, c->srec, c->nbp));
}

// XSPH (or the plain copy of the post-RK velocity) into the sorted set
static void launch_velocity(gfsph_ctx *c, SortedSet &s) {
  const int bsz = c->gbs[2];
  const int *stp = c->cases ? c->stopr : nullptr;
  if (c->P.av_vel && c->gteam[2] == 1)
    SKF_TM_DISPATCH(k_xsph_m, 1, xcd_grid((long)c->n, bsz), bsz, 0,
                    c->stream>>>(c->n, c->P, cells_of(c), c->mask, s, c->nw, s.vel, stp, c->srec));
  else if (c->P.av_vel)
    SKF_TM_DISPATCH(k_xsph_m, 2, xcd_grid((long)c->n * 2, bsz), bsz, 0,
                    c->stream>>>(c->n, c->P, cells_of(c), c->mask, s, c->nw, s.vel, stp, c->srec));
  else
    k_copy_vel<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->nw.vel, s.vel);
}

// h_upgrade for sle = 2 and 3 (sle = 1 is part of the RK kernels); a single
// run's bounding box is then taken again from the new smoothing lengths (the
// RK kernels accumulated it from the old ones)
static void launch_hsml(gfsph_ctx *c, SortedSet &s) {
  if (c->P.sle < 2) return;
  const int bsz = c->gbs[2];
  const int *stp = c->cases ? c->stopr : nullptr;
  SKF_DISPATCH(k_hsml_m, xcd_grid((long)c->n, bsz), bsz, 0,
               c->stream>>>(c->n, c->Pr, cells_of(c), c->mask, s, c->nw, stp, c->srec));
  if (c->R == 1)
    k_bbox<<<c->gridN, GF_BLOCK, 0, c->stream>>>(c->n, c->nw.pos, c->nw.hsml, c->nw.flag, c->part);
}

// adaptive_dt_sph before a step from set[src] (nothing with a fixed dt_sph)
static void enqueue_dt(gfsph_ctx *c, int src) {
  const int *stopr = c->cases ? c->stopr : nullptr;
  if (c->adapt == 7) {
    const SortedSet &o = c->set[src];
    k_dt_fs_scatter<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, o.id, c->nw.pos, o.vel, o.rid, c->fs_xv, c->fs_rk);
    k_dt_fs<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->fs_xv, c->fs_rk, stopr, c->dtm);
    k_dt_apply<<<1, 64, 0, c->stream>>>(c->R, c->P.nstages, c->Pr, c->dtm, c->tm, stopr, c->cases ? 0 : 1);
  } else if (c->adapt) {
    const SortedSet &o = c->set[src];
    k_dt_min<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->adapt, c->P.h_inf, c->P.cgra, c->dxg, o.vel, c->nw.h,
                                                      c->nw.hsml, c->nw.flag, o.rid, stopr, c->dtm);
    k_dt_apply<<<1, 64, 0, c->stream>>>(c->R, c->P.nstages, c->Pr, c->dtm, c->tm, stopr, c->cases ? 0 : 1);
  }
  if (c->cases)
    k_case_pre<<<(c->R + 63) / 64, 64, 0, c->stream>>>(c->R, c->P.nstages, c->Pr, c->tm, c->cs, c->stopr,
                                                         c->cctl.time_plot);
}

// after a step into set[dst]: plot-time pixels, stop conditions
static void enqueue_case_end(gfsph_ctx *c, int dst) {
  if (!c->cases) return;
  const int R = c->R, rb = (R + 63) / 64;
  hipStream_t st = c->stream;
  k_case_due<<<rb, 64, 0, st>>>(R, c->cs, c->stopr, c->cctl.time_plot, c->due, c->outc);
  const SortedSet &s = c->set[dst];
  if (c->case_pixels) {
    CurView v;
    v.pos = c->nw.pos;
    v.vel = s.vel;
    v.h = c->nw.h;
    v.mass = s.mass;
    v.hsml = c->nw.hsml;
    v.u = c->nw.u;
    v.uaux = c->nw.uaux;
    v.flag = c->nw.flag;
    v.id = s.id;
    v.rid = s.rid;
    const long np = (long)c->dem.nx * c->dem.ny;
    k_mc_pixel<<<nblk(c->n), GF_BLOCK, 0, st>>>(c->n, c->dem, v, np, c->mcr_acc, c->mcr_cnt, c->due);
  }
  k_case_out<<<nblk(c->n), GF_BLOCK, 0, st>>>(c->n, c->nw.flag, s.rid, c->outc);
  k_case_sum<<<R, GF_BLOCK, 0, st>>>(c->coff, c->start, s.vel, c->nw.flag, c->stopr, c->cs);
  k_case_stop<<<rb, 64, 0, st>>>(R, c->P.nstages, c->Pr, c->tm, c->cs, c->stopr, c->due, c->outc, c->cnt_r, c->cctl);
}

// time_sph = time_sph + dt_sph after each of nsteps fixed-dt steps (8:1617)
static void advance_time(gfsph_ctx *c, int nsteps) {
  if (c->adapt || c->cases) return;
  for (int r = 0; r < c->R; r++)
    for (int k = 0; k < nsteps; k++) c->time_h[r] = c->time_h[r] + c->dt_h[r];
}

// Enqueues one whole step reading set[src] + nw and producing set[dst] + nw.
static int enqueue_step(gfsph_ctx *c, int src, int dst) {
  enqueue_dt(c, src);
  int rc = neighbour_build(c, src, dst);
  if (rc) return rc;
  SortedSet &s = c->set[dst];
  launch_density(c, s);
  if (!c->staged) launch_pforce(c, s);  // staged: once per stage instead
  launch_stages(c, s);
  launch_hsml(c, s);
  launch_velocity(c, s);
  enqueue_case_end(c, dst);
  return GFSPH_OK;
}

// Steps.  The kernel sequence of GRAPH_STEPS steps (an even number, so the
// sequence ends on the parity it started from) and of a single step from
// either parity are captured once in HIP graphs and replayed: launch gaps of
// ~15 separate launches per step dominate small runs.  GFSPH_NO_GRAPH=1
// launches the kernels directly.
#define GRAPH_STEPS 8

static int capture(gfsph_ctx *c, int nsteps, hipGraphExec_t *out) {
  hipGraph_t g;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = GFSPH_OK, cur = c->cur;
  for (int k = 0; k < nsteps && !rc; k++) {
    rc = enqueue_step(c, cur, 1 - cur);
    cur = 1 - cur;
  }
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, GFSPH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, GFSPH_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

static void drop_graphs(gfsph_ctx *c) {
  for (int k = 0; k < 2; k++) {
    if (c->graph[k]) (void)hipGraphExecDestroy(c->graph[k]);
    if (c->graphG[k]) (void)hipGraphExecDestroy(c->graphG[k]);
    c->graph[k] = c->graphG[k] = nullptr;
  }
}

// Direct launches instead of graphs: GFSPH_NO_GRAPH=1, or a rocprofiler-sdk
// tool attached to the process (rocprofv3 sets ROCP_TOOL_LIBRARIES and
// ROCPROFILER_LIBRARY_CTOR).  Under rocprofv3 --kernel-trace this library's
// processes have ended in a host SIGSEGV inside the HIP runtime while the
// step sequence was being captured or its graph launched; the kernels and
// their arguments are the same either way, so kernel timings under a
// profiler describe the graph path too.
static bool graphs_off() {
  const char *ng = getenv("GFSPH_NO_GRAPH");
  if (ng && ng[0] == '1') return true;
  const char *tl = getenv("ROCP_TOOL_LIBRARIES");
  const char *ct = getenv("ROCPROFILER_LIBRARY_CTOR");
  return (tl && tl[0]) || (ct && ct[0]);
}

static int run_steps(gfsph_ctx *c, int nsteps) {
  static const bool no_graph = graphs_off();
  int rc;
  if (no_graph) {
    for (int k = 0; k < nsteps; k++) {
      if ((rc = enqueue_step(c, c->cur, 1 - c->cur))) return rc;
      c->cur = 1 - c->cur;
    }
    HIPCHK(c, hipGetLastError());
    return GFSPH_OK;
  }
  while (nsteps >= GRAPH_STEPS) {
    if (!c->graphG[c->cur] && (rc = capture(c, GRAPH_STEPS, &c->graphG[c->cur]))) return rc;
    HIPCHK(c, hipGraphLaunch(c->graphG[c->cur], c->stream));
    nsteps -= GRAPH_STEPS;
  }
  for (; nsteps > 0; nsteps--) {
    if (!c->graph[c->cur] && (rc = capture(c, 1, &c->graph[c->cur]))) return rc;
    HIPCHK(c, hipGraphLaunch(c->graph[c->cur], c->stream));
    c->cur = 1 - c->cur;
  }
  return GFSPH_OK;
}

extern "C" {

// Shared by gfsph_create (one realisation) and gfsph_batch_create: R
// realisations of counts[r] particles each, stored one after another; cstR
// ([R][111]) gives each realisation's const(:) (NULL: p->cst for all).
static int create_impl(const gfsph_params *p, const gfsph_dem *dem, int32_t R, const int32_t *counts,
                       const double *cstR, int32_t n, const double *x, const double *v, const double *h,
                       const double *mass, const double *hsml, const double *u, const int32_t *itype,
                       gfsph_ctx **out) {
  if (!out || !p || !dem || n <= 0 || !x || !v || !h || !mass || !hsml || !dem->Z || R < 1) return GFSPH_ERR_ARG;
  *out = nullptr;
  gfsph_ctx *c = new gfsph_ctx();
  std::string err;
  std::vector<DevParams> Ph(R);
  int rc = GFSPH_OK;
  bool any_frictional = false;
  for (int r = 0; r < R && !rc; r++) {
    gfsph_params pr = *p;
    if (cstR) memcpy(pr.cst, cstR + 111 * (size_t)r, sizeof(pr.cst));
    rc = derive_params(&pr, Ph[r], err);
    if (!rc && R > 1) {
      if ((Ph[r].nfric == 4) != (Ph[0].nfric == 4)) {
        err = "a batch is either all Bingham (nfric = 4, its Tau0 carried per realisation) or none";
        rc = GFSPH_ERR_UNSUPPORTED;
      } else if (Ph[r].npwp != Ph[0].npwp) {
        err = "the realisations of a batch must share the pore-pressure column size";
        rc = GFSPH_ERR_ARG;
      } else if (Ph[r].cgra != Ph[0].cgra || Ph[r].alphaS != Ph[0].alphaS) {
        err = "the realisations of a batch must share const(1) and const(14)";
        rc = GFSPH_ERR_ARG;
      }
    }
    if (!rc) any_frictional = any_frictional || Ph[r].if_frictional;
  }
  if (!rc && R > 1) {
    long tot = 0;
    for (int r = 0; r < R; r++) {
      if (counts[r] <= 0) {
        err = "every realisation needs at least one particle";
        rc = GFSPH_ERR_ARG;
        break;
      }
      tot += counts[r];
    }
    if (!rc && tot != n) {
      err = "the realisation counts do not add up to n";
      rc = GFSPH_ERR_ARG;
    }
  }
  if (rc) {
    c->err = err;
    *out = c;
    return rc;
  }
  // terrain zones (ic_topo_props = 10): friction zones and erosion zones
  int zkinds = 0;
  if (dem->nzones < 0 || dem->nzones > 32 || (dem->nzones > 0 && !dem->zones)) {
    c->err = "nzones must be 0..32 with a zone array";
    *out = c;
    return GFSPH_ERR_ARG;
  }
  for (int z = 0; z < dem->nzones; z++) {
    const gfsph_zone &Z = dem->zones[z];
    const int geom = (int)(Z.props[0] + 0.1);
    std::string why;
    if (Z.type != 1 && Z.type != 2 && Z.type != 3)
      why = "basal type " + std::to_string(Z.type) + " (flux and mass zones are outside the implemented path)";
    else if (Z.type == 3 && Z.props[7] < -900)
      why = "pore-pressure zone with the drained-bed value (< -900) is outside the implemented path";
    else if (!(geom >= 1 && geom <= 5) && geom != 101)
      why = "geometry " + std::to_string(geom) + " (1-5 and 101 are implemented)";
    else if (geom == 101 && !dem->zone_nodes)
      why = "geometry 101 needs zone_nodes";
    else if (Z.type == 2 && (Z.law < 0 || Z.law > 3))
      why = "erosion law " + std::to_string(Z.law);
    if (why.empty() && Z.type == 2 && (Z.law == 2 || Z.law == 3))
      for (int r = 0; r < R && why.empty(); r++)
        if (Ph[r].denss <= 1.e-3 || Ph[r].densw_e <= 1.e-3)
          why = "erosion law " + std::to_string(Z.law) + " needs const(17) and const(18) (the reference stops here)";
    if (why.empty() && Z.type == 2 && Z.law == 2)
      for (int r = 0; r < R && why.empty(); r++)
        if (!Ph[r].if_frictional) why = "erosion law 2 reads the frictional law's tan(phi): it needs a frictional nfric";
    if (!why.empty()) {
      c->err = "terrain zone " + std::to_string(z + 1) + ": " + why;
      *out = c;
      return why.find("implemented path") != std::string::npos || why.find("are implemented") != std::string::npos
                 ? GFSPH_ERR_UNSUPPORTED
                 : GFSPH_ERR_ARG;
    }
    zkinds |= Z.type == 1 ? 1 : (Z.type == 2 ? 2 : 4);
    if (geom == 1) c->has_zone1 = 1;
  }
  if (zkinds & 2)
    for (int r = 0; r < R; r++) {
      if (Ph[r].nfric == 4) {
        c->err = "erosion with the Bingham law (nfric = 4) is outside the implemented path";
        *out = c;
        return GFSPH_ERR_UNSUPPORTED;
      }
      Ph[r].eros_on = 1;  // ic_erosion = 1 with basal erosion zones (6:1288-1290)
    }
  for (int r = 0; r < R; r++) {
    Ph[r].eros_single = (counts[r] == 1);
    c->eros = c->eros || Ph[r].eros_on;
    c->staged = c->staged || Ph[r].ic_k0 || Ph[r].costheta;
    if (r == 0) c->law = LAW_FRIC5;
    if (Ph[r].nfric != 5 || Ph[r].costheta || Ph[r].if_V2byR || zkinds) c->law = LAW_ANY;
  }
  c->P = Ph[0];
  c->R = R;
  if (dem->nx < 2 || dem->ny < 2 || !(dem->dx > 0) || !(dem->dy > 0)) {
    *out = c;
    return fail(c, GFSPH_ERR_ARG, "DEM needs at least 2x2 nodes and positive spacing");
  }
  if (itype)
    for (int i = 0; i < n; i++)
      if (itype[i] != 1) {
        *out = c;
        return fail(c, GFSPH_ERR_UNSUPPORTED, "only single-phase soil particles (itype = 1) are implemented");
      }
  c->n = n;
  c->check_finite = p->check_finite;
  for (int k = 0; k < 3; k++) c->gteam[k] = n >= TEAM1_MIN_N ? 1 : 2;
  if (const char *e = getenv("GFSPH_GATHER_LANES")) {  // 1 or 2 lanes per particle (same results)
    const int v = atoi(e);
    if (v == 1 || v == 2)
      for (int k = 0; k < 3; k++) c->gteam[k] = v;
  }
  *out = c;
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->counts_h.assign(counts, counts + R);
  if ((rc = alloc_sets(c, n))) return rc;
  if ((c->P.nfric == 4 || c->staged) && (rc = alloc_bingham(c, n))) return rc;
  hipStream_t st = c->stream;
  bool any_nfric3 = false;
  for (int r = 0; r < R; r++) any_nfric3 = any_nfric3 || Ph[r].nfric == 3;
  if (any_nfric3) {
    // Bingham erosion reads (mass - mass0)/mass; mass0 = the masses as set up
    // (6:6398) until gfsph_set_mass0 says otherwise.  A mass <= 1e-6 stops
    // the reference (6:9450-9455).
    for (int i = 0; i < n; i++)
      if (!(mass[i] > 1.e-6)) return fail(c, GFSPH_ERR_ARG, "nfric = 3 needs every particle mass above 1e-6");
    if ((rc = dalloc(c, &c->mass0, n))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->mass0, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
    c->set[0].mass0 = c->set[1].mass0 = c->mass0;
  }

  // DEM
  long np = (long)dem->nx * dem->ny;
  c->dem.nx = dem->nx;
  c->dem.ny = dem->ny;
  c->dem.xmin = dem->xmin;
  c->dem.ymin = dem->ymin;
  c->dem.dx = dem->dx;
  c->dem.dy = dem->dy;
  c->dem.xmax = dem->xmax;
  c->dem.ymax = dem->ymax;
  if ((rc = dalloc(c, &c->zxy, np))) return rc;
  if (any_frictional)
    if ((rc = dalloc(c, &c->top, 6 * np))) return rc;
  if (dem->topol) {
    double *t9;
    if ((rc = dalloc(c, &t9, 9 * np))) return rc;
    HIPCHK(c, hipMemcpyAsync(t9, dem->topol, sizeof(double) * 9 * np, hipMemcpyHostToDevice, st));
    k_topol_split<<<nblk(np), GF_BLOCK, 0, st>>>(np, t9, c->zxy, c->top);
  } else {
    double *Zd;
    int *im = nullptr;
    if ((rc = dalloc(c, &Zd, np))) return rc;
    HIPCHK(c, hipMemcpyAsync(Zd, dem->Z, sizeof(double) * np, hipMemcpyHostToDevice, st));
    if (dem->in_model) {
      if ((rc = dalloc(c, &im, np))) return rc;
      HIPCHK(c, hipMemcpyAsync(im, dem->in_model, sizeof(int) * np, hipMemcpyHostToDevice, st));
    }
    k_fdiff<<<nblk(np), GF_BLOCK, 0, st>>>(dem->nx, dem->ny, dem->dx, dem->dy, Zd, im, c->zxy, c->top);
  }
  c->dem.zxy = c->zxy;
  c->dem.top = c->top;
  if (zkinds) {
    std::vector<DevZone> zh(dem->nzones);
    for (int z = 0; z < dem->nzones; z++) {
      const gfsph_zone &Z = dem->zones[z];
      DevZone &d = zh[z];
      d.type = Z.type;
      d.law = Z.law;
      d.geom = (int)(Z.props[0] + 0.1);
      d.egeom = (int)(Z.props[0] + 0.01);
      for (int k = 0; k < 16; k++) d.p[k] = Z.props[k];
      // theta = thetadeg pi/180 with pi = 4 atan(1) (2:2682, 2:2716)
      const double theta = Z.props[5] * (4. * atan(1.0)) / 180.;
      d.cth = cos(theta);
      d.sth = sin(theta);
    }
    DevZone *zd;
    if ((rc = dalloc(c, &zd, dem->nzones))) return rc;
    HIPCHK(c, hipMemcpyAsync(zd, zh.data(), sizeof(DevZone) * dem->nzones, hipMemcpyHostToDevice, st));
    c->dem.zones = zd;
    c->dem.nzones = dem->nzones;
    c->dem.zkinds = zkinds;
    if (dem->zone_nodes) {
      int *zn;
      if ((rc = dalloc(c, &zn, np))) return rc;
      HIPCHK(c, hipMemcpyAsync(zn, dem->zone_nodes, sizeof(int) * np, hipMemcpyHostToDevice, st));
      c->dem.znodes = zn;
    }
    if (c->has_zone1) {
      if (dem->in_model) {
        int *im;
        if ((rc = dalloc(c, &im, np))) return rc;
        HIPCHK(c, hipMemcpyAsync(im, dem->in_model, sizeof(int) * np, hipMemcpyHostToDevice, st));
        c->dem.inm = im;
      }
      if (!c->eros) {  // the terrain does not change: Z as given
        double *z0;
        if ((rc = dalloc(c, &z0, np))) return rc;
        HIPCHK(c, hipMemcpyAsync(z0, dem->Z, sizeof(double) * np, hipMemcpyHostToDevice, st));
        c->dem.zcur = z0;
      }
    }
  }
  if (c->eros) {
    // topol(1) = Z, topol(14) = 0, topol(15) = the available depth of each
    // realisation (6:1294-1301), sums cleared
    ErosGrid &g = c->eg;
    g.np = np;
    const size_t tot = (size_t)np * R;
    if ((rc = dalloc(c, &g.z, tot)) || (rc = dalloc(c, &g.eroded, tot)) || (rc = dalloc(c, &g.avail, tot)) ||
        (rc = dalloc(c, &g.aux, 2 * tot)))
      return rc;
    HIPCHK(c, hipMemcpyAsync(g.z, dem->Z, sizeof(double) * np, hipMemcpyHostToDevice, st));
    std::vector<double> av(R);
    for (int r = 0; r < R; r++) av[r] = Ph[r].avail0;
    double *avd;
    if ((rc = dalloc(c, &avd, R))) return rc;
    HIPCHK(c, hipMemcpyAsync(avd, av.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
    // without a surface erosion law (const(21) = 0) the available depth is
    // the terrain's topol(15,:) with values below 1e-6 set to 0 (6:1302-1307)
    const double *ea = nullptr;
    std::vector<double> eah;
    if (zkinds & 2) {
      eah.assign(np, 0.0);
      if (dem->eros_avail)
        for (long k = 0; k < np; k++) eah[k] = dem->eros_avail[k] < 1.e-6 ? 0.0 : dem->eros_avail[k];
      double *ead;
      if ((rc = dalloc(c, &ead, np))) return rc;
      HIPCHK(c, hipMemcpyAsync(ead, eah.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
      ea = ead;
    }
    std::vector<int> lawon(R);
    for (int r = 0; r < R; r++) lawon[r] = Ph[r].eros_law > 0;
    int *lawd;
    if ((rc = dalloc(c, &lawd, R))) return rc;
    HIPCHK(c, hipMemcpyAsync(lawd, lawon.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
    k_eros_init<<<(int)((tot + GF_BLOCK - 1) / GF_BLOCK), GF_BLOCK, 0, st>>>((long)tot, np, avd, ea, lawd, g);
    c->dem.zavail = g.avail;
    if (c->has_zone1) c->dem.zcur = g.z;
  }
  for (int r = 0; r < R; r++) Ph[r].zoff = c->eros ? (long)r * np : 0;

  // particles -> set[0] in caller order; the first neighbour build sorts them
  SortedSet &s0 = c->set[0];
  std::vector<int> ids(n), zeros(n, 0);
  for (int i = 0; i < n; i++) ids[i] = i;
  std::vector<double> zd(n, 0.0);
  HIPCHK(c, hipMemcpyAsync(s0.pos, x, sizeof(double2) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.vel, v, sizeof(double2) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.h, h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.mass, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.hsml, hsml, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.u, u ? u : zd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  // a state that is already initialised (init_input = 0) carries its bottom
  // pore pressure in u: the reference keeps u = Uaux(1,:) (7:956-958)
  const double *ua0 = (!p->init_input && (p->icpwp == 3 || p->icpwp == 2) && u) ? u : zd.data();
  HIPCHK(c, hipMemcpyAsync(s0.uaux, ua0, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.flag, zeros.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(s0.id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  // realisation of every particle and the realisations' slot ranges in the caller's layout
  std::vector<int> rid(n, 0);
  std::vector<int2> rg(R);
  if (R == 1) {
    rg[0] = make_int2(0, n);
  } else {
    int o = 0;
    for (int r = 0; r < R; r++) {
      rg[r] = make_int2(o, o + counts[r]);
      for (int k = 0; k < counts[r]; k++) rid[o + k] = r;
      o += counts[r];
    }
  }
  HIPCHK(c, hipMemcpyAsync(s0.rid, rid.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->rng, rg.data(), sizeof(int2) * R, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->Pr, Ph.data(), sizeof(DevParams) * R, hipMemcpyHostToDevice, st));
  c->Ph = Ph;
  const int nb = nblk(n);
  k_domain_flag<<<nb, GF_BLOCK, 0, st>>>(n, c->dem, s0.pos, s0.flag);
  k_copy_to_new<<<nb, GF_BLOCK, 0, st>>>(n, s0, c->nw, c->P.npwp);
  k_bbox<<<c->gridN, GF_BLOCK, 0, st>>>(n, c->nw.pos, c->nw.hsml, c->nw.flag, c->part);
  if ((rc = neighbour_build(c, 0, 1))) return rc;
  c->cur = 1;
  SortedSet &s1 = c->set[1];
  int resum = (p->init_input && p->icpwp != 0 && p->summation_density) ? 1 : 0;
  if (resum) launch_density(c, s1);
  if (p->init_input) k_init_pwp<<<nb, GF_BLOCK, 0, st>>>(n, c->P, c->P.alphaS, p->pwp0rel, resum, c->S, s1);
  if (p->init_input && c->P.npwp)
    k_init_column<<<nb, GF_BLOCK, 0, st>>>(n, c->Pr, p->pwp0_type, p->pwp0rel, p->xirel, s1);
  // the step begins by re-sorting from the "new" set: publish the state there
  k_copy_to_new<<<nb, GF_BLOCK, 0, st>>>(n, s1, c->nw, c->P.npwp);
  k_bbox<<<c->gridN, GF_BLOCK, 0, st>>>(n, c->nw.pos, c->nw.hsml, c->nw.flag, c->part);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  c->time_h.assign(R, 0.0);
  c->dt_h.assign(R, p->dt);
  c->counts_h.assign(counts, counts + R);
  c->dxg = sqrt(dem->dx * dem->dx + dem->dy * dem->dy);  // dxg_mod (6:6566); flang lowers **0.5 to sqrt
  return GFSPH_OK;
}

int gfsph_create(const gfsph_params *p, const gfsph_dem *dem, int32_t n, const double *x, const double *v,
                 const double *h, const double *mass, const double *hsml, const double *u, const int32_t *itype,
                 gfsph_ctx **out) {
  const int32_t one[1] = {n};
  return create_impl(p, dem, 1, one, nullptr, n, x, v, h, mass, hsml, u, itype, out);
}

int gfsph_batch_create(const gfsph_params *p, const gfsph_dem *dem, int32_t nreal, const int32_t *counts,
                       const double *cst, const double *x, const double *v, const double *h, const double *mass,
                       const double *hsml, const double *u, gfsph_ctx **out) {
  if (!out || nreal < 1 || !counts) return GFSPH_ERR_ARG;
  long n = 0;
  for (int r = 0; r < nreal; r++) n += counts[r] > 0 ? counts[r] : 0;
  if (n <= 0 || n > 0x7fffffffL) return GFSPH_ERR_ARG;
  return create_impl(p, dem, nreal, counts, cst, (int32_t)n, x, v, h, mass, hsml, u, nullptr, out);
}

int gfsph_step(gfsph_ctx *c, int32_t nsteps) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  if (nsteps < 0) return fail(c, GFSPH_ERR_ARG, "nsteps < 0");
  int rc = run_steps(c, nsteps);
  if (rc) return rc;
  advance_time(c, nsteps);
  if (c->check_finite) {
    HIPCHK(c, hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
    k_nonfinite<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, cur_view(c), c->bad);
    int bad = 0;
    HIPCHK(c, hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bad) return fail(c, GFSPH_ERR_NONFINITE, "non-finite particle state");
  }
  return GFSPH_OK;
}

int gfsph_sync(gfsph_ctx *c) {
  if (!c) return GFSPH_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

int gfsph_get_state(gfsph_ctx *c, double *x, double *v, double *h, double *u, double *uaux, double *hsml,
                    int32_t *flag) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  const int n = c->n;
  double2 *dx, *dv;
  double *dh, *du, *dua, *dhs;
  int *df;
  // scratch: reuse the "new" set arrays after a full sync
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<void *> tmp;
  auto A = [&](void **p, size_t b) -> hipError_t {
    hipError_t e = hipMalloc(p, b);
    if (e == hipSuccess) tmp.push_back(*p);
    return e;
  };
  hipError_t e = hipSuccess;
  e = A((void **)&dx, sizeof(double2) * n);
  if (!e) e = A((void **)&dv, sizeof(double2) * n);
  if (!e) e = A((void **)&dh, sizeof(double) * n);
  if (!e) e = A((void **)&du, sizeof(double) * n);
  if (!e) e = A((void **)&dua, sizeof(double) * n);
  if (!e) e = A((void **)&dhs, sizeof(double) * n);
  if (!e) e = A((void **)&df, sizeof(int) * n);
  if (!e) {
    k_unsort<<<nblk(n), GF_BLOCK, 0, c->stream>>>(n, cur_view(c), dx, dv, dh, du, dua, dhs, df);
    e = hipGetLastError();
  }
  if (!e && x) e = hipMemcpyAsync(x, dx, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && v) e = hipMemcpyAsync(v, dv, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && h) e = hipMemcpyAsync(h, dh, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && u) e = hipMemcpyAsync(u, du, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && uaux) e = hipMemcpyAsync(uaux, dua, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && hsml) e = hipMemcpyAsync(hsml, dhs, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e && flag) e = hipMemcpyAsync(flag, df, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  for (void *q : tmp) (void)hipFree(q);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("get_state: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

int gfsph_get_mass(gfsph_ctx *c, double *mass) {
  if (!c || c->n <= 0 || !mass) return GFSPH_ERR_ARG;
  const int n = c->n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *dm = nullptr;
  hipError_t e = hipMalloc((void **)&dm, sizeof(double) * n);
  if (!e) {
    k_unsort_mass<<<nblk(n), GF_BLOCK, 0, c->stream>>>(n, cur_view(c), dm);
    e = hipGetLastError();
  }
  if (!e) e = hipMemcpyAsync(mass, dm, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (dm) (void)hipFree(dm);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("get_mass: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

static int column_copy(gfsph_ctx *c, double *col, bool up) {
  if (!c || c->n <= 0 || !col) return GFSPH_ERR_ARG;
  if (!c->P.npwp) return fail(c, GFSPH_ERR_ARG, "no pore-pressure column (icpwp /= 2)");
  const int n = c->n, N = c->P.npwp;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *tmp = nullptr;
  const size_t b = sizeof(double) * (size_t)n * N;
  hipError_t e = hipMalloc((void **)&tmp, b);
  SortedSet &s = c->set[c->cur];
  if (!e && up) e = hipMemcpyAsync(tmp, col, b, hipMemcpyHostToDevice, c->stream);
  if (!e) {
    if (up)
      k_col_set<<<nblk(n), GF_BLOCK, 0, c->stream>>>(n, N, s.id, tmp, c->nw.col, s.col, c->nw.uaux, c->nw.u);
    else
      k_col_get<<<nblk(n), GF_BLOCK, 0, c->stream>>>(n, N, s.id, c->nw.col, tmp);
    e = hipGetLastError();
  }
  if (!e && !up) e = hipMemcpyAsync(col, tmp, b, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  if (tmp) (void)hipFree(tmp);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("pwp column: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

int gfsph_get_pwp_column(gfsph_ctx *c, double *col) { return column_copy(c, col, false); }

int gfsph_set_pwp_column(gfsph_ctx *c, const double *col) { return column_copy(c, const_cast<double *>(col), true); }

static int erosion_copy(gfsph_ctx *c, int32_t r, double *z, double *eroded, double *avail, bool up) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  if (!c->eros) return fail(c, GFSPH_ERR_ARG, "erosion is off (const(21) = 0 and const(4) <= 1e-4)");
  if (r < 0 || r >= c->R) return fail(c, GFSPH_ERR_ARG, "realisation index out of range");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const long np = c->eg.np;
  const size_t off = (size_t)np * r, b = sizeof(double) * np;
  double *dev[3] = {c->eg.z + off, c->eg.eroded + off, c->eg.avail + off};
  double *host[3] = {z, eroded, avail};
  for (int k = 0; k < 3; k++) {
    if (!host[k]) continue;
    if (up)
      HIPCHK(c, hipMemcpyAsync(dev[k], host[k], b, hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(c, hipMemcpyAsync(host[k], dev[k], b, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

int gfsph_erosion_read(gfsph_ctx *c, int32_t r, double *z, double *eroded, double *avail) {
  return erosion_copy(c, r, z, eroded, avail, false);
}

int gfsph_erosion_set(gfsph_ctx *c, int32_t r, const double *z, const double *eroded, const double *avail) {
  return erosion_copy(c, r, const_cast<double *>(z), const_cast<double *>(eroded), const_cast<double *>(avail),
                      true);
}

int gfsph_neighbour_pairs(int32_t skf, int32_t n, const double *x, const double *hsml, const int32_t *flag,
                          int64_t cap, int64_t *npairs, int32_t *ij) {
  if (n <= 0 || !x || !hsml || !npairs || skf < 1 || skf > 4 || cap < 0 || (cap > 0 && !ij)) return GFSPH_ERR_ARG;
  gfsph_ctx *c = new gfsph_ctx();
  c->n = n;
  c->P.skf = skf;
  c->P.scale_k = (skf == 1 || skf == 4) ? 2 : 3;
  int rc = GFSPH_OK;
  std::vector<int> ids(n), fl(n, 0);
  for (int i = 0; i < n; i++) {
    ids[i] = i;
    if (flag) fl[i] = flag[i] ? 1 : 0;
  }
  std::vector<double> zd(n, 0.0);
  unsigned long long *dcount = nullptr;
  int *dij = nullptr;
  unsigned long long count = 0;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e) rc = GFSPH_ERR_HIP;
  if (!rc) rc = alloc_sets(c, n);
  if (!rc) rc = dalloc(c, &dcount, 1);
  if (!rc) rc = dalloc(c, &dij, 2 * (cap > 0 ? cap : 1));
  if (!rc) {
    hipStream_t st = c->stream;
    SortedSet &s0 = c->set[0];
    e = hipMemcpyAsync(s0.pos, x, sizeof(double2) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(s0.hsml, hsml, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(s0.flag, fl.data(), sizeof(int) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(s0.id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemsetAsync(s0.rid, 0, sizeof(int) * n, st);
    for (auto arr : {s0.h, s0.mass, s0.u, s0.uaux})
      if (!e) e = hipMemcpyAsync(arr, zd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(s0.vel, zd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemsetAsync(dcount, 0, sizeof(unsigned long long), st);
    if (e) rc = GFSPH_ERR_HIP;
    if (!rc) {
      const int nb = nblk(n);
      k_copy_to_new<<<nb, GF_BLOCK, 0, st>>>(n, s0, c->nw, c->P.npwp);
      k_bbox<<<c->gridN, GF_BLOCK, 0, st>>>(n, c->nw.pos, c->nw.hsml, c->nw.flag, c->part);
      rc = neighbour_build(c, 0, 1);
    }
    if (!rc) {
      k_pairs<<<nblk(n), GF_BLOCK, 0, st>>>(n, skf, c->P.scale_k, cells_of(c), c->set[1], cap, dcount, dij);
      e = hipMemcpyAsync(&count, dcount, sizeof(count), hipMemcpyDeviceToHost, st);
      if (!e) e = hipStreamSynchronize(st);
      if (e) rc = GFSPH_ERR_HIP;
    }
    if (!rc) {
      *npairs = (int64_t)count;
      if ((int64_t)count > cap) {
        rc = GFSPH_ERR_CAPACITY;
      } else if (count > 0) {
        e = hipMemcpy(ij, dij, sizeof(int) * 2 * count, hipMemcpyDeviceToHost);
        if (e) rc = GFSPH_ERR_HIP;
      }
    }
  }
  gfsph_destroy(c);
  return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free neighbour set (gfsph_neighbour_pairs_noalloc): the same
// cell-list build and predicate as gfsph_neighbour_pairs, on caller-owned
// device arrays, with every intermediate buffer carved out of one file-scope
// device array.  Each GPU holds its own copy of the array for the lifetime of
// the process; calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
#define NOALLOC_MAX_N 262144
#define NOALLOC_CELLS (4 * NOALLOC_MAX_N + 4096)
#define NOALLOC_SCAN_BLOCKS ((NOALLOC_CELLS + 2 + SCAN_TILE - 1) / SCAN_TILE)
#define NOALLOC_PARTS ((((NOALLOC_MAX_N + GF_BLOCK - 1) / GF_BLOCK) + 7) / 8 * 8)
#define NA_AL(b) (((size_t)(b) + 255) / 256 * 256)
// byte offsets of the pieces (each 256-aligned)
#define NA_ZERO 0
#define NA_KEY (NA_ZERO + NA_AL(4ull * NOALLOC_MAX_N))
#define NA_RANK (NA_KEY + NA_AL(4ull * NOALLOC_MAX_N))
#define NA_PERM (NA_RANK + NA_AL(4ull * NOALLOC_MAX_N))
#define NA_CNT (NA_PERM + NA_AL(4ull * NOALLOC_MAX_N))
#define NA_START (NA_CNT + NA_AL(4ull * (NOALLOC_CELLS + 2)))
#define NA_BSUM (NA_START + NA_AL(4ull * (NOALLOC_CELLS + 2)))
#define NA_PART (NA_BSUM + NA_AL(4ull * NOALLOC_SCAN_BLOCKS))
#define NA_ACC (NA_PART + NA_AL(sizeof(BBoxAcc) * NOALLOC_PARTS))
#define NA_GRID (NA_ACC + NA_AL(sizeof(BBoxAcc)))
#define NA_COFF (NA_GRID + NA_AL(sizeof(GridInfo)))
#define NA_BYTES (NA_COFF + NA_AL(4 * 2))

// the scratch memory of gfsph_neighbour_pairs_noalloc (about 13 MB)
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

__global__ void k_fill_i32(long n, int v, int *p) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long *p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0ull;
}

// Pairs i < j (caller indices) from the counting sort of the caller's
// unsorted arrays: slot -> caller index through perm.  The predicate and
// its argument order (the smaller index first) are those of k_pairs.
__global__ void k_pairs_perm(int n, int scale_k, const GridInfo *grid, const int *start, const int *perm,
                             const double2 *pos, const double *hsml, const int *flag, long long cap,
                             unsigned long long *count, int *ij) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || flag[i]) return;
  const GridInfo g = *grid;
  const double2 pi_ = pos[i];
  const double hi = hsml[i];
  for_candidates(g, start, pi_, [&](int slot) {
    const int j = perm[slot];
    if (j <= i) return;
    double dx, dy, r, mh;
    if (pair_pred(scale_k, pi_, pos[j], hi, hsml[j], dx, dy, r, mh)) {
      const unsigned long long k = atomicAdd(count, 1ull);
      if ((long long)k < cap) {
        ij[2 * k] = i;
        ij[2 * k + 1] = j;
      }
    }
  });
}

// device address of scratch, per device: taken once, with that device current
static std::mutex noalloc_mu;
static unsigned char *noalloc_ptr[64] = {};
static thread_local std::string noalloc_err;

static int noalloc_fail(int code, const std::string &msg) {
  noalloc_err = msg;
  return code;
}

static int noalloc_scratch(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return noalloc_fail(GFSPH_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return noalloc_fail(GFSPH_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return noalloc_fail(GFSPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    noalloc_ptr[dev] = (unsigned char *)p;
  }
  *out = noalloc_ptr[dev];
  return GFSPH_OK;
}

int gfsph_neighbour_pairs_noalloc(int32_t skf, int32_t n, const double *x, const double *hsml,
                                  const int32_t *flag, int64_t cap, uint64_t *npairs, int32_t *ij, void *stream) {
  if (n <= 0 || !x || !hsml || !npairs || skf < 1 || skf > 4 || cap < 0 || (cap > 0 && !ij))
    return noalloc_fail(GFSPH_ERR_ARG, "gfsph_neighbour_pairs_noalloc: invalid argument");
  if (n > NOALLOC_MAX_N)
    return noalloc_fail(GFSPH_ERR_CAPACITY, "gfsph_neighbour_pairs_noalloc: n = " + std::to_string(n) +
                                                " exceeds its maximum of " + std::to_string(NOALLOC_MAX_N) +
                                                " particles (gfsph_neighbour_pairs has no limit)");
  unsigned char *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  int *zero = (int *)(base + NA_ZERO), *key = (int *)(base + NA_KEY), *rank = (int *)(base + NA_RANK);
  int *perm = (int *)(base + NA_PERM), *cnt = (int *)(base + NA_CNT), *start = (int *)(base + NA_START);
  int *bsum = (int *)(base + NA_BSUM), *coff = (int *)(base + NA_COFF);
  BBoxAcc *part = (BBoxAcc *)(base + NA_PART), *acc = (BBoxAcc *)(base + NA_ACC);
  GridInfo *grid = (GridInfo *)(base + NA_GRID);
  const int scale_k = (skf == 1 || skf == 4) ? 2 : 3;
  const int nb = nblk(n), gridN = (nb + 7) / 8 * 8;
  unsigned long long *count = (unsigned long long *)npairs;
  // the array keeps the last call's contents: clear what must start zero
  k_fill_i32<<<nb, GF_BLOCK, 0, st>>>(n, 0, zero);
  k_fill_i32<<<(int)((NOALLOC_CELLS + 2 + GF_BLOCK - 1) / GF_BLOCK), GF_BLOCK, 0, st>>>(NOALLOC_CELLS + 2, 0, cnt);
  k_zero_u64<<<1, 64, 0, st>>>(count);
  const int *fl = flag ? flag : zero;
  k_bbox<<<gridN, GF_BLOCK, 0, st>>>(n, (const double2 *)x, hsml, fl, part);
  k_grid<<<1, GF_BLOCK, 0, st>>>(gridN, part, acc, grid, coff, scale_k, NOALLOC_CELLS);
  k_keys<<<nb, GF_BLOCK, 0, st>>>(n, (const double2 *)x, fl, zero, grid, coff, coff + 1, key, rank, cnt);
  k_scan_reduce<<<NOALLOC_SCAN_BLOCKS, GF_BLOCK, 0, st>>>(coff + 1, cnt, bsum);
  k_scan_down<<<NOALLOC_SCAN_BLOCKS, GF_BLOCK, 0, st>>>(coff + 1, cnt, bsum, start, n);
  k_scatter<<<nb, GF_BLOCK, 0, st>>>(n, key, rank, start, perm);
  k_pairs_perm<<<nb, GF_BLOCK, 0, st>>>(n, scale_k, grid, start, perm, (const double2 *)x, hsml, fl, cap, count,
                                        ij);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return noalloc_fail(GFSPH_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

int gfsph_raster_idw(gfsph_ctx *c, double *out, double *wsum) {
  if (!c || !out || !wsum) return GFSPH_ERR_ARG;
  long np = (long)c->dem.nx * c->dem.ny;
  int rc;
  if (!c->raster5 && (rc = dalloc(c, &c->raster5, 5 * np + 5 * np))) return rc;
  double *acc5 = c->raster5, *nrm = c->raster5 + 5 * np;
  HIPCHK(c, hipMemsetAsync(acc5, 0, sizeof(double) * 5 * np, c->stream));
  k_raster_idw<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->dem, cur_view(c), acc5, -1);
  k_raster_norm<<<nblk(np), GF_BLOCK, 0, c->stream>>>(np, acc5, nrm, nrm + 4 * np);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, nrm, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(wsum, nrm + 4 * np, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

int gfsph_mc_pixel_accumulate(gfsph_ctx *c) {
  if (!c) return GFSPH_ERR_ARG;
  long np = (long)c->dem.nx * c->dem.ny;
  int rc;
  if (!c->mc_acc) {
    if ((rc = dalloc(c, &c->mc_acc, 4 * np)) || (rc = dalloc(c, &c->mc_cnt, np))) return rc;
    HIPCHK(c, hipMemsetAsync(c->mc_acc, 0, sizeof(double) * 4 * np, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mc_cnt, 0, sizeof(int) * np, c->stream));
  }
  k_mc_pixel<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->dem, cur_view(c), 0, c->mc_acc, c->mc_cnt, nullptr);
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

int gfsph_mc_pixel_read(gfsph_ctx *c, double *acc, int32_t *cnt, int32_t reset) {
  if (!c || !acc || !cnt) return GFSPH_ERR_ARG;
  long np = (long)c->dem.nx * c->dem.ny;
  if (!c->mc_acc) {
    memset(acc, 0, sizeof(double) * 4 * np);
    memset(cnt, 0, sizeof(int) * np);
    return GFSPH_OK;
  }
  HIPCHK(c, hipMemcpyAsync(acc, c->mc_acc, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(cnt, c->mc_cnt, sizeof(int) * np, hipMemcpyDeviceToHost, c->stream));
  if (reset) {
    HIPCHK(c, hipMemsetAsync(c->mc_acc, 0, sizeof(double) * 4 * np, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mc_cnt, 0, sizeof(int) * np, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

int gfsph_profile_step(gfsph_ctx *c, int32_t nrep, int32_t cap, int32_t *nkernels, const char **names, double *ms) {
  if (!c || nrep <= 0 || !nkernels) return GFSPH_ERR_ARG;
  if (c->cases) return fail(c, GFSPH_ERR_UNSUPPORTED, "per-kernel profiling runs without case control");
  static const char *kn[] = {"neighbour_build", "k_density", "k_pforce", "k_rk", "k_xsph", "step_total"};
  const int K = 6;
  *nkernels = K;
  if (cap < K || !names || !ms) return GFSPH_ERR_CAPACITY;
  hipEvent_t ev[6];
  for (int k = 0; k < 6; k++) HIPCHK(c, hipEventCreate(&ev[k]));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < nrep; r++) {
    hipStream_t st = c->stream;
    int src = c->cur, dst = 1 - c->cur;
    enqueue_dt(c, src);
    advance_time(c, 1);
    HIPCHK(c, hipEventRecord(ev[0], st));
    int rc = neighbour_build(c, src, dst);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(ev[1], st));
    SortedSet &s = c->set[dst];
    launch_density(c, s);
    HIPCHK(c, hipEventRecord(ev[2], st));
    launch_pforce(c, s);
    HIPCHK(c, hipEventRecord(ev[3], st));
    launch_stages(c, s);
    HIPCHK(c, hipEventRecord(ev[4], st));
    launch_velocity(c, s);
    HIPCHK(c, hipEventRecord(ev[5], st));
    c->cur = dst;
    HIPCHK(c, hipEventSynchronize(ev[5]));
    float t;
    for (int k = 0; k < 5; k++) {
      HIPCHK(c, hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      acc[k] += t;
    }
    HIPCHK(c, hipEventElapsedTime(&t, ev[0], ev[5]));
    acc[5] += t;
  }
  for (int k = 0; k < 6; k++) (void)hipEventDestroy(ev[k]);
  for (int k = 0; k < K; k++) {
    names[k] = kn[k];
    ms[k] = acc[k] / nrep;
  }
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

static int ensure_mcr(gfsph_ctx *c) {
  if (c->mcr_acc) return GFSPH_OK;
  const long np = (long)c->dem.nx * c->dem.ny;
  int rc;
  if ((rc = dalloc(c, &c->mcr_acc, 4 * np * c->R)) || (rc = dalloc(c, &c->mcr_cnt, np * c->R))) return rc;
  HIPCHK(c, hipMemsetAsync(c->mcr_acc, 0, sizeof(double) * 4 * np * c->R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->mcr_cnt, 0, sizeof(int) * np * c->R, c->stream));
  return GFSPH_OK;
}

int gfsph_batch_pixel_accumulate(gfsph_ctx *c) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  int rc = ensure_mcr(c);
  if (rc) return rc;
  const long np = (long)c->dem.nx * c->dem.ny;
  k_mc_pixel<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->dem, cur_view(c), np, c->mcr_acc, c->mcr_cnt,
                                                      nullptr);
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

int gfsph_batch_pixel_read(gfsph_ctx *c, int32_t r, double *acc, int32_t *cnt) {
  if (!c || r < 0 || r >= c->R) return GFSPH_ERR_ARG;
  int rc = ensure_mcr(c);
  if (rc) return rc;
  const long np = (long)c->dem.nx * c->dem.ny;
  if (acc)
    HIPCHK(c, hipMemcpyAsync(acc, c->mcr_acc + 4 * np * r, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, c->stream));
  if (cnt) HIPCHK(c, hipMemcpyAsync(cnt, c->mcr_cnt + np * r, sizeof(int) * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

int gfsph_batch_hazard(gfsph_ctx *c, double hthr, double *stats, int32_t *nprob, double *maxs, double *pexc) {
  if (!c || !stats || !nprob || !maxs || !pexc) return GFSPH_ERR_ARG;
  int rc = ensure_mcr(c);
  if (rc) return rc;
  const long np = (long)c->dem.nx * c->dem.ny;
  // one temporary block: stats [np][8], maxs [np][2], pexc [np], nprob [np]
  void *tmp = nullptr;
  const size_t bytes = sizeof(double) * 11 * np + sizeof(int) * np;
  HIPCHK(c, hipMalloc(&tmp, bytes));
  double *d_stats = (double *)tmp, *d_maxs = d_stats + 8 * np, *d_pexc = d_maxs + 2 * np;
  int *d_nprob = (int *)(d_pexc + np);
  k_mc_hazard<<<nblk(np), GF_BLOCK, 0, c->stream>>>(np, c->R, c->mcr_acc, c->mcr_cnt, hthr, d_stats, d_nprob, d_maxs,
                                                     d_pexc);
  hipError_t e = hipGetLastError();
  if (!e) e = hipMemcpyAsync(stats, d_stats, sizeof(double) * 8 * np, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipMemcpyAsync(nprob, d_nprob, sizeof(int) * np, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipMemcpyAsync(maxs, d_maxs, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipMemcpyAsync(pexc, d_pexc, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream);
  if (!e) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("batch hazard: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

// Work-group size of each gather kernel, timed on the context's current
// sorted set: the three gathers are idempotent there (they recompute the same
// depths, masks, forces and XSPH velocities), so tuning does not advance or
// change the state.  Needs at least one completed step.
int gfsph_autotune(gfsph_ctx *c, int32_t reps, int32_t *chosen) {
  if (!c || c->n <= 0 || reps <= 0) return GFSPH_ERR_ARG;
  static const int sizes[] = {64, 128, 256, 512};
  SortedSet &s = c->set[c->cur];
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int best[3], best_tm[3];
  for (int k = 0; k < 3; k++) {
    float tbest = 1e30f;
    best[k] = c->gbs[k];
    best_tm[k] = c->gteam[k];
    for (int tm = 2; tm >= 1; tm--) {  // lanes per particle: the same bits either way
      for (int bsz : sizes) {
        if (k == 2 && !c->P.av_vel) break;
        c->gbs[k] = bsz;
        c->gteam[k] = tm;
        for (int w = 0; w < 2; w++) {  // warm-up, then timed
          HIPCHK(c, hipEventRecord(e0, c->stream));
          for (int r = 0; r < reps; r++) {
            if (k == 0) launch_density(c, s);
            else if (k == 1) launch_pforce(c, s);
            else launch_velocity(c, s);
          }
          HIPCHK(c, hipEventRecord(e1, c->stream));
        }
        HIPCHK(c, hipEventSynchronize(e1));
        float t;
        HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
        if (t < tbest) {
          tbest = t;
          best[k] = bsz;
          best_tm[k] = tm;
        }
      }
    }
    c->gbs[k] = best[k];
    c->gteam[k] = best_tm[k];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(c, hipGetLastError());
  drop_graphs(c);  // recaptured with the chosen sizes
  if (chosen)
    for (int k = 0; k < 3; k++) chosen[k] = best[k];
  return GFSPH_OK;
}

// ---------------------------------------------------------------------------
// adjoint host side
// ---------------------------------------------------------------------------
// A checkpoint holds what the next step reads: of the sorted set the masses,
// ids, velocities and realisations (the neighbour build takes the rest from
// the step output), of the step output the positions, depths, smoothing
// lengths, pore pressures and flags; the Bingham carry; and what the next
// neighbour build takes its grids from -- a batch's slot ranges of the saved
// sorted set (a single run's bounding box is recomputed from the restored
// step output at restore: min / max reductions, so the grid is the one the
// forward built).  84 B per particle.
static int alloc_snap(gfsph_ctx *c, gfsph_ctx::Snap &sn) {
  const int n = c->n;
  int rc;
  sn.s = SortedSet{};
  sn.o = NewSet{};
  if ((rc = dalloc(c, &sn.s.mass, n)) || (rc = dalloc(c, &sn.s.id, n)) || (rc = dalloc(c, &sn.s.vel, n)) ||
      (rc = dalloc(c, &sn.s.rid, n)) || (rc = dalloc(c, &sn.o.pos, n)) || (rc = dalloc(c, &sn.o.h, n)) ||
      (rc = dalloc(c, &sn.o.hsml, n)) || (rc = dalloc(c, &sn.o.u, n)) || (rc = dalloc(c, &sn.o.uaux, n)) ||
      (rc = dalloc(c, &sn.o.flag, n)))
    return rc;
  sn.carry = nullptr;
  sn.rng = nullptr;
  if (c->P.nfric == 4 && (rc = dalloc(c, &sn.carry, c->R))) return rc;
  if (c->R > 1 && (rc = dalloc(c, &sn.rng, c->R))) return rc;
  return GFSPH_OK;
}

#define DCOPY(dst, src, count) \
  HIPCHK(c, hipMemcpyAsync((dst), (src), sizeof(*(src)) * (size_t)(count), hipMemcpyDeviceToDevice, c->stream))

static int snap_save(gfsph_ctx *c, gfsph_ctx::Snap &sn) {
  const int n = c->n;
  const SortedSet &s = c->set[c->cur];
  sn.cur = c->cur;
  DCOPY(sn.s.mass, s.mass, n);
  DCOPY(sn.s.id, s.id, n);
  DCOPY(sn.s.vel, s.vel, n);
  DCOPY(sn.s.rid, s.rid, n);
  DCOPY(sn.o.pos, c->nw.pos, n);
  DCOPY(sn.o.h, c->nw.h, n);
  DCOPY(sn.o.hsml, c->nw.hsml, n);
  DCOPY(sn.o.u, c->nw.u, n);
  DCOPY(sn.o.uaux, c->nw.uaux, n);
  DCOPY(sn.o.flag, c->nw.flag, n);
  if (sn.carry) DCOPY(sn.carry, c->bs.carry, c->R);
  if (sn.rng) {
    if (c->first_build)
      DCOPY(sn.rng, c->rng, c->R);
    else
      k_rng<<<(c->R + GF_BLOCK - 1) / GF_BLOCK, GF_BLOCK, 0, c->stream>>>(c->R, c->coff, c->start, sn.rng);
  }
  return GFSPH_OK;
}

static int snap_restore(gfsph_ctx *c, gfsph_ctx::Snap &sn) {
  const int n = c->n;
  c->cur = sn.cur;
  SortedSet &s = c->set[c->cur];
  DCOPY(s.mass, sn.s.mass, n);
  DCOPY(s.id, sn.s.id, n);
  DCOPY(s.vel, sn.s.vel, n);
  DCOPY(s.rid, sn.s.rid, n);
  DCOPY(c->nw.pos, sn.o.pos, n);
  DCOPY(c->nw.h, sn.o.h, n);
  DCOPY(c->nw.hsml, sn.o.hsml, n);
  DCOPY(c->nw.u, sn.o.u, n);
  DCOPY(c->nw.uaux, sn.o.uaux, n);
  DCOPY(c->nw.flag, sn.o.flag, n);
  if (sn.carry) DCOPY(c->bs.carry, sn.carry, c->R);
  if (sn.rng) {
    DCOPY(c->rng, sn.rng, c->R);
    c->first_build = true;  // the next build takes these ranges as they are
  } else {
    k_bbox<<<c->gridN, GF_BLOCK, 0, c->stream>>>(c->n, c->nw.pos, c->nw.hsml, c->nw.flag, c->part);
  }
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

// copies cell starts [0, *ntot + 2) to dst when they fit its cap entries, else
// raises *over (the tape is then recorded again with larger buffers)
__global__ void k_copy_starts(const int *ntot, const int *src, int *dst, int cap, int *over) {
  const int need = *ntot + 2;
  if (need > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(over, need);
    return;
  }
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < need; k += gridDim.x * blockDim.x) dst[k] = src[k];
}

__global__ void k_scatter_vel(int n, const int *id, const double2 *vel, double2 *vn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vn[id[i]] = vel[i];
}

static int alloc_tape(gfsph_ctx *c, TapeRec &t) {
  const int n = c->n;
  int rc;
  if ((rc = dalloc(c, &t.nb, n)) || (rc = dalloc(c, &t.id, n)) || (rc = dalloc(c, &t.rid, n)) ||
      (rc = dalloc(c, &t.sflag, n)) || (rc = dalloc(c, &t.oflag, n)) || (rc = dalloc(c, &t.S, n)) ||
      (rc = dalloc(c, &t.ain, n)) || (rc = dalloc(c, &t.vp, n)) || (rc = dalloc(c, &t.vn, n)) ||
      (rc = dalloc(c, &t.mask, NB_WORDS * (size_t)n)) || (rc = dalloc(c, &t.grid, c->R)) ||
      (rc = dalloc(c, &t.coff, c->R + 1)))
    return rc;
  // the cell starts at the size the forward pass used (its largest cell count
  // with a margin; grown if a recomputed step needs more)
  t.start_cap = std::min(c->cap_cells + 2, std::max(4096, c->ad_cells + c->ad_cells / 2 + 2));
  if ((rc = dalloc(c, &t.start, (size_t)t.start_cap))) return rc;
  t.bt = nullptr;
  if (c->P.nfric == 4 && (rc = dalloc(c, &t.bt, 2 * (size_t)c->P.nstages * n))) return rc;
  return GFSPH_OK;
}

// one recomputed step with its record
static int taped_step(gfsph_ctx *c, TapeRec &t) {
  const int n = c->n;
  const SortedSet &s0 = c->set[c->cur];
  k_scatter_vel<<<nblk(n), GF_BLOCK, 0, c->stream>>>(n, s0.id, s0.vel, t.vn);
  int rc;
  if (t.bt) {  // Bingham: the stages' Tau0 are recorded from inside the step, launched directly
    c->tape_bt = t.bt;
    rc = enqueue_step(c, c->cur, 1 - c->cur);
    c->tape_bt = nullptr;
    if (rc) return rc;
    c->cur = 1 - c->cur;
  } else if ((rc = run_steps(c, 1))) {
    return rc;
  }
  const SortedSet &s = c->set[c->cur];
  DCOPY(t.nb, s.nb, n);
  DCOPY(t.id, s.id, n);
  DCOPY(t.rid, s.rid, n);
  DCOPY(t.sflag, s.flag, n);
  DCOPY(t.oflag, c->nw.flag, n);
  DCOPY(t.S, c->S, n);
  DCOPY(t.ain, c->ain, n);
  DCOPY(t.vp, c->nw.vel, n);
  DCOPY(t.mask, c->mask, NB_WORDS * (size_t)n);
  k_copy_starts<<<std::max(1, std::min(1024, (t.start_cap + GF_BLOCK - 1) / GF_BLOCK)), GF_BLOCK, 0, c->stream>>>(
      c->coff + c->R, c->start, t.start, t.start_cap, c->ad_over);
  DCOPY(t.grid, c->grid, c->R);
  DCOPY(t.coff, c->coff, c->R + 1);
  return GFSPH_OK;
}

static int adjoint_supported(gfsph_ctx *c) {
  if (c->adapt) return fail(c, GFSPH_ERR_UNSUPPORTED, "the adjoint runs with a fixed dt_sph (ic_adapt_dt_sph = 0)");
  if (c->cases) return fail(c, GFSPH_ERR_UNSUPPORTED, "the adjoint runs without case control");
  if (c->eros) return fail(c, GFSPH_ERR_UNSUPPORTED, "the adjoint runs without erosion");
  if (c->staged) return fail(c, GFSPH_ERR_UNSUPPORTED, "the adjoint runs without ic_K0 and ic_cosTheta");
  if (c->dem.nzones) return fail(c, GFSPH_ERR_UNSUPPORTED, "the adjoint runs without terrain zones");
  for (const DevParams &P : c->Ph) {
    if (P.if_V2byR || P.icpwp != 0 || P.sle != 0 || P.skf != 1 || P.pa_sph != 2 || P.nfric != c->Ph[0].nfric)
      return fail(c, GFSPH_ERR_UNSUPPORTED,
                  "the adjoint covers pa_sph = 2, skf = 1, icpwp = 0, sle = 0 and one law for every realisation");
    if (!((P.nfric == 5 && P.xi_voellmy >= 0.0) || (P.nfric == 4 && P.nstages <= BING_MAX_STAGES)))
      return fail(c, GFSPH_ERR_UNSUPPORTED,
                  "the adjoint covers nfric = 5 (frictional / Voellmy, xi >= 0) and nfric = 4 (Bingham)");
  }
  return GFSPH_OK;
}

static int cells_now(gfsph_ctx *c, int *out) {
  HIPCHK(c, hipMemcpyAsync(out, c->coff + c->R, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFSPH_OK;
}

static void dfree(gfsph_ctx *c, void *p) {
  for (size_t k = 0; k < c->allocs.size(); k++)
    if (c->allocs[k] == p) {
      c->allocs.erase(c->allocs.begin() + (long)k);
      break;
    }
  (void)hipFree(p);
}

int gfsph_adjoint_set_tape(gfsph_ctx *c, int32_t tape_steps) {
  if (!c || tape_steps < 0) return GFSPH_ERR_ARG;
  c->ad_tape = tape_steps;
  return GFSPH_OK;
}

int gfsph_adjoint_forward(gfsph_ctx *c, int32_t nsteps, int32_t ckpt_every) {
  if (!c || c->n <= 0 || nsteps <= 0 || ckpt_every <= 0) return GFSPH_ERR_ARG;
  int rc = adjoint_supported(c);
  if (rc) return rc;
  const int nseg = (nsteps + ckpt_every - 1) / ckpt_every;
  if ((int)c->snaps.size() < nseg + 1) {
    const size_t old = c->snaps.size();
    c->snaps.resize(nseg + 1);
    for (size_t k = old; k < c->snaps.size(); k++)
      if ((rc = alloc_snap(c, c->snaps[k]))) return rc;
  }
  if (!c->ad_over && (rc = dalloc(c, &c->ad_over, 1))) return rc;
  c->ad_ckpt = ckpt_every;
  c->ad_nsteps = nsteps;
  int cells = 0;
  for (int seg = 0; seg < nseg; seg++) {
    if ((rc = snap_save(c, c->snaps[seg]))) return rc;
    const int L = std::min(ckpt_every, nsteps - seg * ckpt_every);
    if ((rc = run_steps(c, L))) return rc;
    if ((rc = cells_now(c, &cells))) return rc;
    c->ad_cells = std::max(c->ad_cells, cells);
  }
  if ((rc = snap_save(c, c->snaps[nseg]))) return rc;
  HIPCHK(c, hipGetLastError());
  return GFSPH_OK;
}

int gfsph_adjoint_backward(gfsph_ctx *c, const double *lam_x, const double *lam_v, const double *lam_h,
                           double *g_mu, double *g_xi, double *g_mass, double *g_x0, double *g_v0) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  if (c->ad_nsteps <= 0) return fail(c, GFSPH_ERR_ARG, "gfsph_adjoint_forward has not run");
  const int n = c->n, K = c->ad_ckpt, N = c->ad_nsteps, nseg = (N + K - 1) / K, R = c->R;
  const bool bing = c->P.nfric == 4;
  const int S = c->P.nstages;
  hipStream_t st = c->stream;
  int rc;
  AdjWork &w = c->aw;
  if (!c->LX) {
    if ((rc = dalloc(c, &c->LX, n)) || (rc = dalloc(c, &c->LV, n)) || (rc = dalloc(c, &c->LM, n)) ||
        (rc = dalloc(c, &c->LH, n)) || (rc = dalloc(c, &c->GP, n)) || (rc = dalloc(c, &c->GQ, n)) ||
        (rc = dalloc(c, &w.lxp, n)) || (rc = dalloc(c, &w.lvpp, n)) || (rc = dalloc(c, &w.lvp, n)) ||
        (rc = dalloc(c, &w.la, n)) || (rc = dalloc(c, &w.lx, n)) || (rc = dalloc(c, &w.lxn, n)) ||
        (rc = dalloc(c, &w.lvn, n)) || (rc = dalloc(c, &w.lS, n)) || (rc = dalloc(c, &w.lm, n)) ||
        (rc = dalloc(c, &w.act, n)))
      return rc;
    if (bing && ((rc = dalloc(c, &c->lamTau, n)) || (rc = dalloc(c, &c->lamCand, (size_t)S * n)) ||
                 (rc = dalloc(c, &c->tlead, c->bs.ntiles)) || (rc = dalloc(c, &c->tcin, c->bs.ntiles)) ||
                 (rc = dalloc(c, &c->thas, c->bs.ntiles)) || (rc = dalloc(c, &c->bcar, R))))
      return rc;
  }
  std::vector<double> zero(2 * (size_t)n, 0.0);
  HIPCHK(c, hipMemcpyAsync(c->LX, lam_x ? lam_x : zero.data(), sizeof(double2) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->LV, lam_v ? lam_v : zero.data(), sizeof(double2) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->LH, lam_h ? lam_h : zero.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemsetAsync(c->LM, 0, sizeof(double) * n, st));
  HIPCHK(c, hipMemsetAsync(c->GP, 0, sizeof(double) * n, st));
  HIPCHK(c, hipMemsetAsync(c->GQ, 0, sizeof(double) * n, st));
  if (bing) HIPCHK(c, hipMemsetAsync(c->bcar, 0, sizeof(double) * R, st));
  const int nb = nblk(n);
  const BinghamState &b = c->bs;
  // at most T steps are recorded at once: a longer segment is split, from
  // its checkpoint, into at most ADJ_BRANCH pieces with checkpoints of their
  // own, recursively, each piece swept back in reverse order
  const int T = (c->ad_tape > 0 && c->ad_tape < K) ? c->ad_tape : K;
  if ((int)c->tape.size() < T) {
    const size_t old = c->tape.size();
    c->tape.resize(T);
    for (size_t k = old; k < c->tape.size(); k++)
      if ((rc = alloc_tape(c, c->tape[k]))) return rc;
  }
  bool first_back = true;  // the step swept back first (the last step) takes lam_h
  auto step_back = [&](TapeRec &tr) {
    k_adj_load<<<nb, GF_BLOCK, 0, st>>>(n, c->P.npoin0, tr, c->LX, c->LV, first_back ? c->LH : nullptr, w);
    first_back = false;
    if (c->P.av_vel)
      k_adj_xsph<<<xcd_grid((long)n * ADJ_TM, GF_BLOCK), GF_BLOCK, 0, st>>>(n, c->P, tr, w);
    else
      (void)hipMemcpyAsync(w.lvp, w.lvpp, sizeof(double2) * n, hipMemcpyDeviceToDevice, st);
    if (bing) {
      for (int sd = S - 1; sd >= 0; sd--) {
        k_adj_bing_tau<<<nb, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, tr, w, sd, c->lamCand, c->lamTau);
        const double *cand = tr.bt + (size_t)(S + sd) * n;
        k_seg_reduce<<<b.ntiles, GF_BLOCK, 0, st>>>(b.tiles, c->lamTau, cand, c->tlead, c->thas);
        k_seg_top<<<(R + 63) / 64, 64, 0, st>>>(R, b.rtile, c->tlead, c->thas, c->tcin, c->bcar);
        k_seg_down<<<b.ntiles, GF_BLOCK, 0, st>>>(b.tiles, c->lamTau, cand, c->tcin, c->lamCand + (size_t)sd * n);
      }
      k_adj_stage_bing<<<nb, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, tr, w, c->lamCand, c->GP, c->GQ);
    } else {
      k_adj_stage<<<nb, GF_BLOCK, 0, st>>>(n, c->Pr, c->dem, tr, w, c->GP, c->GQ);
    }
    k_adj_pairs<<<xcd_grid((long)n * ADJ_TM, GF_BLOCK), GF_BLOCK, 0, st>>>(n, c->P, tr, w, c->LX, c->LV, c->LM);
  };
  // records the L <= T steps from checkpoint sn and sweeps them back
  auto record_and_back = [&](gfsph_ctx::Snap &sn, int L) -> int {
    int r;
    if ((r = snap_restore(c, sn))) return r;
    HIPCHK(c, hipMemsetAsync(c->ad_over, 0, sizeof(int), st));
    for (int t = 0; t < L; t++)
      if ((r = taped_step(c, c->tape[t]))) return r;
    int over = 0;
    HIPCHK(c, hipMemcpyAsync(&over, c->ad_over, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (over > 0) {  // a recorded step had more cells than the buffers: grow them and record again
      for (TapeRec &tr : c->tape) {
        dfree(c, tr.start);
        tr.start = nullptr;
        tr.start_cap = std::min(c->cap_cells + 2, over + over / 2);
        if ((r = dalloc(c, &tr.start, (size_t)tr.start_cap))) return r;
      }
      if ((r = snap_restore(c, sn))) return r;
      for (int t = 0; t < L; t++)
        if ((r = taped_step(c, c->tape[t]))) return r;
    }
    for (int t = L - 1; t >= 0; t--) step_back(c->tape[t]);
    HIPCHK(c, hipGetLastError());
    return GFSPH_OK;
  };
  std::function<int(gfsph_ctx::Snap &, int, int)> sweep = [&](gfsph_ctx::Snap &sn, int L, int depth) -> int {
    if (L <= T) return record_and_back(sn, L);
    const int sub = (L + ADJ_BRANCH - 1) / ADJ_BRANCH, np = (L + sub - 1) / sub;
    int r;
    if ((int)c->snapsL.size() <= depth) c->snapsL.resize(depth + 1);
    std::vector<gfsph_ctx::Snap> &lv = c->snapsL[depth];
    if ((int)lv.size() < np) {
      const size_t old = lv.size();
      lv.resize(np);
      for (size_t k = old; k < lv.size(); k++)
        if ((r = alloc_snap(c, lv[k]))) return r;
    }
    if ((r = snap_restore(c, sn))) return r;
    for (int j = 0; j < np; j++) {
      if ((r = snap_save(c, lv[j]))) return r;
      if ((r = run_steps(c, std::min(sub, L - j * sub)))) return r;
    }
    for (int j = np - 1; j >= 0; j--)
      if ((r = sweep(c->snapsL[depth][j], std::min(sub, L - j * sub), depth + 1))) return r;
    return GFSPH_OK;
  };
  for (int seg = nseg - 1; seg >= 0; seg--)
    if ((rc = sweep(c->snaps[seg], std::min(K, N - seg * K), 0))) return rc;
  // back to the state the forward pass ended in
  if ((rc = snap_restore(c, c->snaps[nseg]))) return rc;
  std::vector<double> gp(n), gq(n);
  HIPCHK(c, hipMemcpyAsync(gp.data(), c->GP, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipMemcpyAsync(gq.data(), c->GQ, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  if (g_mass) HIPCHK(c, hipMemcpyAsync(g_mass, c->LM, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  if (g_x0) HIPCHK(c, hipMemcpyAsync(g_x0, c->LX, sizeof(double2) * n, hipMemcpyDeviceToHost, st));
  if (g_v0) HIPCHK(c, hipMemcpyAsync(g_v0, c->LV, sizeof(double2) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  // each realisation's parameter cotangents: its particles' sums, in id order
  int first = 0;
  for (int r = 0; r < R; r++) {
    const int nr = R == 1 ? n : c->counts_h[r];
    double sp = 0.0, sq = 0.0;
    for (int k = first; k < first + nr; k++) {
      sp += gp[k];
      sq += gq[k];
    }
    if (g_mu) g_mu[r] = sp;
    if (g_xi) g_xi[r] = sq;
    first += nr;
  }
  return GFSPH_OK;
}

int gfsph_raster_misfit(gfsph_ctx *c, const double *target, double *loss, double *lam_x, double *lam_h) {
  if (!c || c->n <= 0 || !target || !loss) return GFSPH_ERR_ARG;
  const long np = (long)c->dem.nx * c->dem.ny;
  const int n = c->n, R = c->R;
  hipStream_t st = c->stream;
  int rc;
  if (!c->raster5 && (rc = dalloc(c, &c->raster5, 5 * np + 5 * np))) return rc;
  const int nbn = nblk(np);
  void *tmp = nullptr;
  const size_t bytes = sizeof(double) * (np + 2 * np + (size_t)nbn * R) + sizeof(double2) * n + sizeof(double) * n;
  HIPCHK(c, hipMalloc(&tmp, bytes));
  double *d_t = (double *)tmp, *d_res = d_t + np, *d_part = d_res + 2 * np;
  double2 *d_lx = (double2 *)(d_part + (size_t)nbn * R);
  double *d_lh = (double *)(d_lx + n);
  std::vector<double> part((size_t)nbn * R);
  hipError_t e = hipMemcpyAsync(d_t, target, sizeof(double) * np, hipMemcpyHostToDevice, st);
  // each realisation against the same target, one after the other on the one raster
  for (int r = 0; r < R && !e; r++) {
    const int rs = R == 1 ? -1 : r;
    e = hipMemsetAsync(c->raster5, 0, sizeof(double) * 5 * np, st);
    if (!e) {
      k_raster_idw<<<nblk(n), GF_BLOCK, 0, st>>>(n, c->dem, cur_view(c), c->raster5, rs);
      k_misfit_nodes<<<nbn, GF_BLOCK, 0, st>>>(np, c->raster5, d_t, d_res, d_part + (size_t)nbn * r);
      k_misfit_vjp<<<nblk(n), GF_BLOCK, 0, st>>>(n, c->dem, cur_view(c), c->raster5, d_res, d_lx, d_lh, rs);
      e = hipGetLastError();
    }
  }
  if (!e) e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * nbn * R, hipMemcpyDeviceToHost, st);
  if (!e && lam_x) e = hipMemcpyAsync(lam_x, d_lx, sizeof(double2) * n, hipMemcpyDeviceToHost, st);
  if (!e && lam_h) e = hipMemcpyAsync(lam_h, d_lh, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (!e) e = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("raster misfit: ") + hipGetErrorString(e));
  for (int r = 0; r < R; r++) {
    double L = 0.0;
    for (int k = 0; k < nbn; k++) L += part[(size_t)nbn * r + k];
    loss[r] = L;
  }
  return GFSPH_OK;
}

int gfsph_counts(gfsph_ctx *c, int32_t *n, int32_t *nactive) {
  if (!c) return GFSPH_ERR_ARG;
  if (n) *n = c->n;
  if (nactive) {
    HIPCHK(c, hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
    k_count_active<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, c->nw.flag, c->bad);
    int na = 0;
    HIPCHK(c, hipMemcpyAsync(&na, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *nactive = (int32_t)na;
  }
  return GFSPH_OK;
}

// time_sph and dt_sph live on the device while the time step adapts or the
// case control runs (they then change inside the step graphs), else on the host
static bool device_time(const gfsph_ctx *c) { return c->adapt || c->cases; }

static int time_to_device(gfsph_ctx *c) {
  const int R = c->R;
  int rc;
  if (!c->tm && ((rc = dalloc(c, &c->tm, R)) || (rc = dalloc(c, &c->dtm, R)))) return rc;
  std::vector<unsigned long long> init(R, DT_INIT);
  HIPCHK(c, hipMemcpy(c->tm, c->time_h.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->dtm, init.data(), sizeof(unsigned long long) * R, hipMemcpyHostToDevice));
  return GFSPH_OK;
}

static int time_to_host(gfsph_ctx *c) {
  const int R = c->R;
  std::vector<DevParams> ph(R);
  HIPCHK(c, hipMemcpy(c->time_h.data(), c->tm, sizeof(double) * R, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ph.data(), c->Pr, sizeof(DevParams) * R, hipMemcpyDeviceToHost));
  for (int r = 0; r < R; r++) c->dt_h[r] = ph[r].dt;
  c->P.dt = ph[0].dt;
  c->P.fin = ph[0].fin;
  return GFSPH_OK;
}

int gfsph_set_adaptive_dt(gfsph_ctx *c, int32_t mode) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  // mode 3 halves its running minimum at every pair of the pair list
  // (6:6587-6590): with more than ~1,100 pairs it underflows to 0 and the
  // reference's dt_sph becomes 0, below that it depends on the list order
  if (mode != 0 && mode != 1 && mode != 2 && mode != 4 && mode != 7 && mode != 10)
    return fail(c, GFSPH_ERR_UNSUPPORTED,
                "ic_adapt_dt_sph must be 0, 1, 2, 4, 7 or 10 (3 drives dt_sph to 0 on any flow of more than ~1,100 "
                "pairs)");
  if (mode == c->adapt) return GFSPH_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc;
  if (mode == 7 && !c->fs_xv && ((rc = dalloc(c, &c->fs_xv, c->n)) || (rc = dalloc(c, &c->fs_rk, c->n)))) return rc;
  const bool before = device_time(c);
  if (!before && mode && (rc = time_to_device(c))) return rc;
  if (before && !mode && !c->cases && (rc = time_to_host(c))) return rc;
  drop_graphs(c);
  c->adapt = mode;
  return GFSPH_OK;
}

int gfsph_batch_scale_pwp(gfsph_ctx *c, const double *factor) {
  if (!c || c->n <= 0 || !factor) return GFSPH_ERR_ARG;
  if (!c->P.icpwp) return fail(c, GFSPH_ERR_ARG, "no pore pressure (icpwp = 0)");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *f = nullptr;
  hipError_t e = hipMalloc((void **)&f, sizeof(double) * c->R);
  if (!e) e = hipMemcpy(f, factor, sizeof(double) * c->R, hipMemcpyHostToDevice);
  if (!e) {
    k_scale_pwp<<<nblk(c->n), GF_BLOCK, 0, c->stream>>>(c->n, f, c->set[c->cur], c->nw, c->P.npwp);
    e = hipGetLastError();
  }
  if (!e) e = hipStreamSynchronize(c->stream);
  if (f) (void)hipFree(f);
  if (e) return fail(c, GFSPH_ERR_HIP, std::string("scale_pwp: ") + hipGetErrorString(e));
  return GFSPH_OK;
}

int gfsph_set_mass0(gfsph_ctx *c, const double *mass0) {
  if (!c || c->n <= 0 || !mass0) return GFSPH_ERR_ARG;
  if (!c->mass0) return fail(c, GFSPH_ERR_ARG, "mass0 is read by nfric = 3 only");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->mass0, mass0, sizeof(double) * c->n, hipMemcpyHostToDevice));
  return GFSPH_OK;
}

int gfsph_set_time(gfsph_ctx *c, double time) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->time_h.assign(c->R, time);
  if (device_time(c)) HIPCHK(c, hipMemcpy(c->tm, c->time_h.data(), sizeof(double) * c->R, hipMemcpyHostToDevice));
  if (c->cases) {
    std::vector<CaseState> h(c->R);
    HIPCHK(c, hipMemcpy(h.data(), c->cs, sizeof(CaseState) * c->R, hipMemcpyDeviceToHost));
    for (auto &q : h) q.tglob = time;
    HIPCHK(c, hipMemcpy(c->cs, h.data(), sizeof(CaseState) * c->R, hipMemcpyHostToDevice));
  }
  return GFSPH_OK;
}

int gfsph_time(gfsph_ctx *c, double *time, double *dt) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  const int R = c->R;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!device_time(c)) {
    if (time) memcpy(time, c->time_h.data(), sizeof(double) * R);
    if (dt) memcpy(dt, c->dt_h.data(), sizeof(double) * R);
    return GFSPH_OK;
  }
  if (time) HIPCHK(c, hipMemcpy(time, c->tm, sizeof(double) * R, hipMemcpyDeviceToHost));
  if (dt) {
    std::vector<DevParams> ph(R);
    HIPCHK(c, hipMemcpy(ph.data(), c->Pr, sizeof(DevParams) * R, hipMemcpyDeviceToHost));
    for (int r = 0; r < R; r++) dt[r] = ph[r].dt;
  }
  return GFSPH_OK;
}

int gfsph_batch_cases(gfsph_ctx *c, const gfsph_case_ctl *ctl) {
  if (!c || c->n <= 0 || !ctl) return GFSPH_ERR_ARG;
  if (c->P.nfric == 4) return fail(c, GFSPH_ERR_UNSUPPORTED, "case control does not cover Bingham (nfric = 4)");
  if (!(ctl->time_plot >= 0.0) || !(ctl->time_end > 0.0))
    return fail(c, GFSPH_ERR_ARG, "time_plot must be >= 0 and time_end > 0");
  const int R = c->R;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc;
  if (!c->cs && ((rc = dalloc(c, &c->cs, R)) || (rc = dalloc(c, &c->stopr, R)) || (rc = dalloc(c, &c->due, R)) ||
                 (rc = dalloc(c, &c->outc, R)) || (rc = dalloc(c, &c->cnt_r, R))))
    return rc;
  if (ctl->pixels && (rc = ensure_mcr(c))) return rc;
  if (!device_time(c) && (rc = time_to_device(c))) return rc;
  // loop state of a fresh case (8:1459-1462, 8:104-106): dt_max = dt is
  // not used by Driver_Flow_23; t_plot_reset = 0; time = time_sph
  std::vector<double> t0(R);
  HIPCHK(c, hipMemcpy(t0.data(), c->tm, sizeof(double) * R, hipMemcpyDeviceToHost));
  std::vector<CaseState> h(R);
  std::vector<int> zero(R, 0);
  for (int r = 0; r < R; r++) {
    h[r] = CaseState{};
    h[r].tglob = t0[r];
  }
  HIPCHK(c, hipMemcpy(c->cs, h.data(), sizeof(CaseState) * R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->stopr, zero.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->due, zero.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->outc, zero.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->cnt_r, c->counts_h.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  c->cctl = CaseCtl{ctl->vstop, ctl->timestop, ctl->dtimestop, ctl->dt, ctl->time_end, ctl->time_plot};
  c->case_pixels = ctl->pixels ? 1 : 0;
  c->cases = 1;
  drop_graphs(c);
  return GFSPH_OK;
}

int gfsph_batch_status(gfsph_ctx *c, int32_t *stopped, int32_t *nsteps, int32_t *nstopped) {
  if (!c || c->n <= 0) return GFSPH_ERR_ARG;
  if (!c->cases) return fail(c, GFSPH_ERR_ARG, "case control is not enabled (gfsph_batch_cases)");
  const int R = c->R;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> st(R);
  HIPCHK(c, hipMemcpy(st.data(), c->stopr, sizeof(int) * R, hipMemcpyDeviceToHost));
  if (stopped) memcpy(stopped, st.data(), sizeof(int) * R);
  if (nsteps) {
    std::vector<CaseState> h(R);
    HIPCHK(c, hipMemcpy(h.data(), c->cs, sizeof(CaseState) * R, hipMemcpyDeviceToHost));
    for (int r = 0; r < R; r++) nsteps[r] = h[r].nsteps;
  }
  if (nstopped) {
    int k = 0;
    for (int v : st) k += v ? 1 : 0;
    *nstopped = k;
  }
  return GFSPH_OK;
}

const char *gfsph_last_error(gfsph_ctx *c) { return c ? c->err.c_str() : noalloc_err.c_str(); }

void gfsph_destroy(gfsph_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
