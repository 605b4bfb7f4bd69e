/*
 * gfsph_oracle.c -- serial FP64 CPU restatement of the Geoflow-SPH
 * depth-integrated SPH propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check the
 * GPU's results or to time the CPU baseline; the HIP product path never calls
 * it.
 *
 * Every routine cites the reference file:line it follows (file prefixes as in
 * "Codes Geoflow_SPH/<n> ...f90").  Parity pins (see tests/test_oracle_*.py):
 *   neighbour set        - reference dumps (bit-exact set), brute force,
 *                          lattice closed form
 *   kernels skf 1..4     - closed forms, unit integral, antisymmetry,
 *                          reference dumps of w / dwdx
 *   density / pressure   - reference dumps, uniform-lattice closed form,
 *                          momentum invariant
 *   rheology + limiter   - reference dumps (state after 100 RK4 steps),
 *                          inclined-plane closed forms (frictional, Voellmy
 *                          terminal speed), Bingham root identity; laws 22
 *                          and 23 (m = 1, 2, 1/3) reference dumps; law 3
 *                          (Bingham erosion) reference dumps while the eroded
 *                          fraction is <= 0.01, beyond it each Bingham regime
 *                          equals a reference-pinned law bit for bit (the
 *                          reference reads an unset local there)
 *   pore pressure        - reference dumps, exponential-decay closed form
 *   RK4 / RK2            - reference dumps (RK4), ODE order tests (RK2 has no
 *                          reference counterpart: pinned by order + closed form)
 *   XSPH / sle=1,2,3     - reference dumps
 *   adaptive dt          - reference dumps of modes 2 and 10 (states and
 *                          time_sph after 1, 10, 100 steps); modes 1 and 4
 *                          share their terms (pinned through mode 10)
 *   terrain zones        - Which_Topo_Zone at 20,909 points (reference
 *                          dump), reference runs with friction and erosion
 *                          zones (states, masses, terrain), each zone shown
 *                          to change the reference's state
 *   rasters              - single-particle closed forms, constant-field
 *                          reproduction (no reference dump: the reference
 *                          writes them only as GiD text)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gfsph_oracle.h"

/* ------------------------------------------------------------------ */
/* Parameters derived the way Input_SW / ExtForces_SW derive them.      */
/* ------------------------------------------------------------------ */

typedef struct {
  int nfric, if_frictional, if_V2byR, ic_cosTheta;
  double cgra, dens, xi_voellmy, constMu_CF, K_Bagnold;
  double tauy0, constK, visco, tanfi8, hfrict0, Bfact, comp, alphaS;
  double hrelpw2; /* const(14)*const(14) evaluated in REAL(4): 6:8470 */
  double alphaS2; /* same product in Get_Pwp_SW: 6:10934              */
  double c73;
  double fact_dens, fact_pwp_max, fact_pwp_min, fpwp_max, fpwp_min;
  int scale_k;
  double pi;
  /* erosion (Input_SW 6:1180-1184, 6:1288-1308; Get_Erosion_Rate_SW
     6:10289-10510; ExtForces_SW 6:8664-8693, 6:8724-8812) */
  int eros_on;   /* ic_erosion: const(21) > 1e-9 (after the const(4) form)  */
  int eros_law;  /* const(21) + 0.1: 1 Hungr, 2 slope-angle, 3 basal stress */
  double keros, eros_c23, eros_c24, eros_c25, denss, densw_e, dens_bar_eff, slip_fact, avail0;
  double Cv; /* consolidation coefficient 4 Bfact / pi^2 (6:10926) */
  /* terrain zones: the current erosion state [npoig][3] (topol(1), (14),
     (15)) while a run erodes, else NULL */
  const double *eros;
} derived_t;

static void derive(const gfo_params *p, derived_t *d) {
  const double *c = p->cst;
  memset(d, 0, sizeof(*d));
  d->pi = 4.0 * atan(1.0); /* 6:792 (FP64 under -fdefault-real-8) */
  d->nfric = (int)(c[5] + 0.01); /* 6:8538 */
  d->if_frictional = 0;
  d->if_V2byR = 0;
  if ((d->nfric >= 5 && d->nfric <= 10) || d->nfric == 26 || d->nfric == 27) {
    d->if_frictional = 1;
    if (d->nfric == 7 || d->nfric == 8 || d->nfric == 10 || d->nfric == 26 || d->nfric == 27)
      d->if_V2byR = 1;
  }
  d->ic_cosTheta = (int)(c[16] + 0.01);
  d->cgra = c[1];
  d->dens = c[2];
  d->xi_voellmy = c[3]; /* 6:8421 */
  if (d->xi_voellmy < 0) d->constMu_CF = -(25. * c[3] / 4.); /* 6:8449 */
  d->K_Bagnold = c[3];
  d->tauy0 = c[6];
  d->constK = c[7];
  d->visco = c[8];
  d->tanfi8 = c[9];
  d->hfrict0 = c[10];
  d->Bfact = c[13];
  d->comp = c[15];
  d->alphaS = c[14];
  d->hrelpw2 = (double)((float)c[14] * (float)c[14]);
  d->alphaS2 = d->hrelpw2;
  d->c73 = 7. / 3.;
  /* DFs_FSs_dens_and_pwp_factors, ic_fct_dens = 0 branch (6:8998-9008) */
  d->fact_dens = 1.0;
  d->fact_pwp_max = 1.0;
  d->fact_pwp_min = 0.0;
  d->fpwp_max = 1.0;
  d->fpwp_min = 0.0;
  if (p->skf == 1 || p->skf == 4) d->scale_k = 2; /* 5a:850-856 */
  else d->scale_k = 3;
  /* erosion constants: const(4) > 1e-4 is the older form of law 1 (6:1180-1184) */
  double c21 = c[21], c22 = c[22];
  if (c[4] > 1.e-4) {
    c21 = 1;
    c22 = c[4];
  }
  d->eros_on = c21 > 1.e-9;
  d->eros_law = (int)(c21 + 0.1);
  d->keros = c22;
  d->eros_c23 = c[23];
  d->eros_c24 = c[24];
  d->eros_c25 = c[25];
  d->denss = c[17];
  d->densw_e = c[18];
  /* DFs_FSs_dens_and_pwp_factors, ic_fct_dens = 0 (6:9001-9008) */
  d->dens_bar_eff = d->dens - (c[18] < 1.e-3 ? 1000.0 : c[18]);
  /* nconst = 110 > 45: Vbottom = alpha Vtop (6:8672-8679) */
  d->slip_fact = 1.0;
  if (c[47] > 1.e-6) d->slip_fact = 2 * c[47] / (1. + c[47]);
  /* available erosion depth topol(15,:) (6:1294-1301) */
  d->avail0 = (c[48] > 1.e-6) ? c[48] : 999.;
  d->Cv = 4 * c[13] / (d->pi * d->pi);
}

/* ------------------------------------------------------------------ */
/* a2. kernel_TOOLS (5a:2062-2180), ndimn = 2                           */
/* ------------------------------------------------------------------ */

void gfo_kernel(int skf, double r, double dx, double dy, double hsml, double *w, double *dwx,
                double *dwy) {
  const double pi = 4.0 * atan(1.0);
  double q = r / hsml, factor, a;
  *w = 0.0;
  *dwx = 0.0;
  *dwy = 0.0;
  if (skf == 1) {
    factor = 15.e0 / (7.e0 * pi * hsml * hsml);
    if (q >= 0 && q <= 1.e0) {
      *w = factor * (2. / 3. - q * q + q * q * q / 2.);
      a = factor * (-2. + 3. / 2. * q) / (hsml * hsml);
      *dwx = a * dx;
      *dwy = a * dy;
    } else if (q > 1.e0 && q <= 2) {
      *w = factor * 1.e0 / 6.e0 * ((2. - q) * (2. - q) * (2. - q));
      a = -(factor * 1.e0 / 6.e0 * 3. * ((2. - q) * (2. - q)) / hsml);
      *dwx = a * (dx / r);
      *dwy = a * (dy / r);
    }
  } else if (skf == 2) {
    factor = 1.e0 / (hsml * hsml * pow(pi, 2 / 2.));
    if (q >= 0 && q <= 3) {
      *w = factor * exp(-q * q);
      *dwx = *w * (-2. * dx / hsml / hsml);
      *dwy = *w * (-2. * dy / hsml / hsml);
    }
  } else if (skf == 3) {
    factor = 7.e0 / (478.e0 * pi * hsml * hsml);
    double t3 = 3 - q, t2 = 2 - q, t1 = 1 - q;
    double t3_4 = t3 * t3 * t3 * t3, t2_4 = t2 * t2 * t2 * t2;
    if (q >= 0 && q <= 1) {
      *w = factor * (t3_4 * t3 - 6 * t2_4 * t2 + 15 * (t1 * t1 * t1 * t1 * t1));
      a = factor * ((-120 + 120 * q - 50 * q * q) / (hsml * hsml));
      *dwx = a * dx;
      *dwy = a * dy;
    } else if (q > 1 && q <= 2) {
      *w = factor * (t3_4 * t3 - 6 * t2_4 * t2);
      a = factor * (-5 * t3_4 + 30 * t2_4) / hsml;
      *dwx = a * (dx / r);
      *dwy = a * (dy / r);
    } else if (q > 2 && q <= 3) {
      *w = factor * (t3_4 * t3);
      a = factor * (-5 * t3_4) / hsml;
      *dwx = a * (dx / r);
      *dwy = a * (dy / r);
    }
  } else if (skf == 4) {
    factor = 7.e0 / (4.e0 * pi * hsml * hsml);
    if (q >= 0 && q <= 2) {
      double b = 1 - q / 2;
      *w = factor * ((b * b * b * b) * (2 * q + 1));
      a = factor * (-5 * (b * b * b)) * ((1 / hsml / hsml));
      *dwx = a * dx;
      *dwy = a * dy;
    }
  }
}

/* ------------------------------------------------------------------ */
/* a1. neighbour pairs: predicate of grid_find_NEW_tools3 (5a:1010-1019) */
/*     r = sqrt(dx*dx+dy*dy) < scale_k * ((hsml_i+hsml_j)/2)            */
/* Candidate pairs come from an axis-aligned cell list whose cell is     */
/* 2*max(hsml) (5a:912) times scale_k/2, so every pair passing the       */
/* predicate lies in adjacent cells.  Output pairs have i < j (0-based), */
/* ordered by i then j, with w / dwdx evaluated at dx = x_i - x_j.       */
/* ------------------------------------------------------------------ */

static inline int pair_test(int scale_k, double xi, double yi, double xj, double yj, double hi,
                            double hj, double *dx, double *dy, double *r, double *mh) {
  double ddx = xi - xj, ddy = yi - yj;
  double dr = ddx * ddx;
  double t = ddy * ddy; /* built with -ffp-contract=off: no FMA */
  dr = dr + t;
  *dx = ddx;
  *dy = ddy;
  *r = sqrt(dr);
  *mh = (hi + hj) / 2.;
  return *r < scale_k * (*mh);
}

typedef struct {
  int64_t n, cap;
  int32_t *i, *j;
} pairvec;

static void pv_push(pairvec *v, int32_t a, int32_t b) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 1024;
    v->i = (int32_t *)realloc(v->i, sizeof(int32_t) * v->cap);
    v->j = (int32_t *)realloc(v->j, sizeof(int32_t) * v->cap);
  }
  v->i[v->n] = a;
  v->j[v->n] = b;
  v->n++;
}

static int cmp_pair(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* Builds the sorted (i<j) pair list. brute != 0 -> O(N^2) direct_find (7:1827) */
static void build_pairs(int n, int skf, const double *x, const double *hsml, const int *flag,
                        int brute, pairvec *out) {
  int scale_k = (skf == 1 || skf == 4) ? 2 : 3;
  double dx, dy, r, mh;
  out->n = 0;
  if (brute) {
    for (int i = 0; i < n; i++) {
      if (flag[i]) continue;
      for (int j = i + 1; j < n; j++) {
        if (flag[j]) continue;
        if (pair_test(scale_k, x[2 * i], x[2 * i + 1], x[2 * j], x[2 * j + 1], hsml[i], hsml[j],
                      &dx, &dy, &r, &mh))
          pv_push(out, i, j);
      }
    }
    return;
  }
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300, hmax = 0;
  int nact = 0;
  for (int i = 0; i < n; i++) {
    if (flag[i]) continue;
    nact++;
    if (x[2 * i] < xmin) xmin = x[2 * i];
    if (x[2 * i] > xmax) xmax = x[2 * i];
    if (x[2 * i + 1] < ymin) ymin = x[2 * i + 1];
    if (x[2 * i + 1] > ymax) ymax = x[2 * i + 1];
    if (hsml[i] > hmax) hmax = hsml[i];
  }
  if (nact == 0) return;
  double cell = hmax * scale_k * 1.000001 + 1e-12;
  int ncx = (int)((xmax - xmin) / cell) + 1, ncy = (int)((ymax - ymin) / cell) + 1;
  int64_t ncell = (int64_t)ncx * ncy;
  int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * n);
  int *cx = (int *)malloc(sizeof(int) * n), *cy = (int *)malloc(sizeof(int) * n);
  for (int64_t c = 0; c < ncell; c++) head[c] = -1;
  for (int i = n - 1; i >= 0; i--) {
    if (flag[i]) continue;
    int a = (int)((x[2 * i] - xmin) / cell), b = (int)((x[2 * i + 1] - ymin) / cell);
    if (a >= ncx) a = ncx - 1;
    if (b >= ncy) b = ncy - 1;
    cx[i] = a;
    cy[i] = b;
    next[i] = head[(int64_t)b * ncx + a];
    head[(int64_t)b * ncx + a] = i;
  }
  for (int i = 0; i < n; i++) {
    if (flag[i]) continue;
    for (int b = cy[i] - 1; b <= cy[i] + 1; b++) {
      if (b < 0 || b >= ncy) continue;
      for (int a = cx[i] - 1; a <= cx[i] + 1; a++) {
        if (a < 0 || a >= ncx) continue;
        for (int j = head[(int64_t)b * ncx + a]; j >= 0; j = next[j]) {
          if (j <= i) continue;
          if (pair_test(scale_k, x[2 * i], x[2 * i + 1], x[2 * j], x[2 * j + 1], hsml[i], hsml[j],
                        &dx, &dy, &r, &mh))
            pv_push(out, i, j);
        }
      }
    }
  }
  free(head);
  free(next);
  free(cx);
  free(cy);
  /* canonical order (i, j) ascending */
  int64_t *key = (int64_t *)malloc(sizeof(int64_t) * (out->n ? out->n : 1));
  for (int64_t k = 0; k < out->n; k++) key[k] = ((int64_t)out->i[k] << 32) | (uint32_t)out->j[k];
  qsort(key, out->n, sizeof(int64_t), cmp_pair);
  for (int64_t k = 0; k < out->n; k++) {
    out->i[k] = (int32_t)(key[k] >> 32);
    out->j[k] = (int32_t)(key[k] & 0xffffffff);
  }
  free(key);
}

int64_t gfo_pairs(int n, int skf, const double *x, const double *hsml, const int32_t *flag,
                  int brute, int64_t cap, int32_t *ij, double *w, double *dwdx) {
  pairvec v = {0, 0, NULL, NULL};
  build_pairs(n, skf, x, hsml, flag, brute, &v);
  int scale_k = (skf == 1 || skf == 4) ? 2 : 3;
  if (cap >= v.n) {
    for (int64_t k = 0; k < v.n; k++) {
      int i = v.i[k], j = v.j[k];
      double dx, dy, r, mh, ww, gx, gy;
      pair_test(scale_k, x[2 * i], x[2 * i + 1], x[2 * j], x[2 * j + 1], hsml[i], hsml[j], &dx,
                &dy, &r, &mh);
      gfo_kernel(skf, r, dx, dy, mh, &ww, &gx, &gy);
      ij[2 * k] = i;
      ij[2 * k + 1] = j;
      if (w) w[k] = ww;
      if (dwdx) {
        dwdx[2 * k] = gx;
        dwdx[2 * k + 1] = gy;
      }
    }
  }
  int64_t np = v.n;
  free(v.i);
  free(v.j);
  return np;
}

/* ------------------------------------------------------------------ */
/* a7. Get_Z_Topo (2:2395-2491): bilinear lookup of nodal fields.      */
/* topol layout [npoig][9]: Z, Zx, Zy, top4..top9                       */
/* ------------------------------------------------------------------ */

typedef struct {
  double Zp, Zx, Zy, t4, t5, t6, t7, t8, t9;
} topo_pt;

static void get_z_topo(const gfo_dem *d, double xx, double yy, int want_top, topo_pt *o) {
  int ipx = (int)((xx - d->xmin) / d->dx + 1);
  int ipy = (int)((yy - d->ymin) / d->dy + 1);
  double xmax = d->xmin + (d->nx - 1) * d->dx; /* reference keeps Xmaxg from the file: 2:247 */
  double ymax = d->ymin + (d->ny - 1) * d->dy;
  if (d->xmax_set) {
    xmax = d->xmax;
    ymax = d->ymax;
  }
  if (ipx == d->nx && fabs(xmax - xx) <= 1.e-3) ipx = ipx - 1;
  if (ipy == d->ny && fabs(ymax - yy) <= 1.e-3) ipy = ipy - 1;
  /* guard (the reference warns and reads past the grid, 2:2421-2435) */
  if (ipx < 1) ipx = 1;
  if (ipy < 1) ipy = 1;
  if (ipx > d->nx - 1) ipx = d->nx - 1;
  if (ipy > d->ny - 1) ipy = d->ny - 1;
  int64_t ipoig = (int64_t)(ipy - 1) * d->nx + ipx; /* 1-based */
  int64_t n1 = ipoig - 1, n2 = ipoig, n3 = ipoig + d->nx, n4 = n3 - 1; /* 0-based */
  double xref = d->xmin + (ipx - 1) * d->dx;
  double yref = d->ymin + (ipy - 1) * d->dy;
  double xi = (xx - xref) / d->dx, eta = (yy - yref) / d->dy;
  double sh1 = (1. - xi) * (1. - eta), sh2 = xi * (1. - eta), sh3 = xi * eta,
         sh4 = (1. - xi) * eta;
  const double *T = d->topol;
#define TV(k, n) T[(n)*9 + (k)]
  o->Zp = sh1 * TV(0, n1) + sh2 * TV(0, n2) + sh3 * TV(0, n3) + sh4 * TV(0, n4);
  if (d->in_model) {
    int kp = d->in_model[n1] + d->in_model[n2] + d->in_model[n3] + d->in_model[n4];
    if (kp < 4) {
      double z = d->in_model[n1] * sh1 * TV(0, n1) + d->in_model[n2] * sh2 * TV(0, n2) +
                 d->in_model[n3] * sh3 * TV(0, n3) + d->in_model[n4] * sh4 * TV(0, n4);
      double den = d->in_model[n1] * sh1 + d->in_model[n2] * sh2 + d->in_model[n3] * sh3 +
                   d->in_model[n4] * sh4;
      o->Zp = z;
      if (den > 1.e-3) o->Zp = z / den;
    }
  }
  o->Zx = sh1 * TV(1, n1) + sh2 * TV(1, n2) + sh3 * TV(1, n3) + sh4 * TV(1, n4);
  o->Zy = sh1 * TV(2, n1) + sh2 * TV(2, n2) + sh3 * TV(2, n3) + sh4 * TV(2, n4);
  if (want_top) {
    double *tt[6] = {&o->t4, &o->t5, &o->t6, &o->t7, &o->t8, &o->t9};
    for (int k = 0; k < 6; k++)
      *tt[k] = sh1 * TV(3 + k, n1) + sh2 * TV(3 + k, n2) + sh3 * TV(3 + k, n3) +
               sh4 * TV(3 + k, n4);
  }
#undef TV
}

/* Get_Z_Topo's Z (2:2437-2462) from the current Z (zc[k * zs], node k) */
static double z_topo_cur(const gfo_dem *d, const double *zc, int zs, double xx, double yy) {
  int ipx = (int)((xx - d->xmin) / d->dx + 1);
  int ipy = (int)((yy - d->ymin) / d->dy + 1);
  double xmax = d->xmin + (d->nx - 1) * d->dx, ymax = d->ymin + (d->ny - 1) * d->dy;
  if (d->xmax_set) {
    xmax = d->xmax;
    ymax = d->ymax;
  }
  if (ipx == d->nx && fabs(xmax - xx) <= 1.e-3) ipx = ipx - 1;
  if (ipy == d->ny && fabs(ymax - yy) <= 1.e-3) ipy = ipy - 1;
  if (ipx < 1) ipx = 1;
  if (ipy < 1) ipy = 1;
  if (ipx > d->nx - 1) ipx = d->nx - 1;
  if (ipy > d->ny - 1) ipy = d->ny - 1;
  int64_t n1 = (int64_t)(ipy - 1) * d->nx + ipx - 1, n2 = n1 + 1, n3 = n1 + 1 + d->nx, n4 = n3 - 1;
  double xi = (xx - (d->xmin + (ipx - 1) * d->dx)) / d->dx, eta = (yy - (d->ymin + (ipy - 1) * d->dy)) / d->dy;
  double sh1 = (1. - xi) * (1. - eta), sh2 = xi * (1. - eta), sh3 = xi * eta, sh4 = (1. - xi) * eta;
  const double z1 = zc[n1 * zs], z2 = zc[n2 * zs], z3 = zc[n3 * zs], z4 = zc[n4 * zs];
  double Zp = sh1 * z1 + sh2 * z2 + sh3 * z3 + sh4 * z4;
  if (d->in_model) {
    const int32_t *m = d->in_model;
    if (m[n1] + m[n2] + m[n3] + m[n4] < 4) {
      Zp = m[n1] * sh1 * z1 + m[n2] * sh2 * z2 + m[n3] * sh3 * z3 + m[n4] * sh4 * z4;
      double den = m[n1] * sh1 + m[n2] * sh2 + m[n3] * sh3 + m[n4] * sh4;
      if (den > 1.e-3) Zp = Zp / den;
    }
  }
  return Zp;
}

/* Which_Topo_Zone (2:2635-2808): the first zone of basal type btype whose
   region holds (xx, yy), or -1.  typ_geom = int(props(1) + 0.1); geometry
   1-3 a range of Z, x, y (ZZdiscr = (max - v)(v - min) >= 0), 4 a
   rectangle, 5 a rectangle turned props(6) degrees with pi = 4 atan(1),
   101 the nodes n1, n1+1, n1+1+nx, n1+nx of the point's cell (the cell
   index as Get_Z_Topo forms it) against the zone's node list.  zc, zs: the
   current Z of node k at zc[k * zs], for geometry 1. */
static int which_zone(const gfo_dem *d, const double *zc, int zs, double xx, double yy, int btype) {
  for (int z = 0; z < d->nzones; z++) {
    const gfo_zone *Z = &d->zones[z];
    if (Z->type != btype) continue;
    const int geom = (int)(Z->props[0] + 0.1);
    const double *P = Z->props;
    double discr = -999;
    if (geom <= 3) {
      double Zp = geom == 1 ? z_topo_cur(d, zc, zs, xx, yy) : (geom == 2 ? xx : yy);
      discr = (P[2] - Zp) * (Zp - P[1]);
    } else if (geom == 4) {
      double XX = (P[2] - xx) * (xx - P[1]), YY = (P[4] - yy) * (yy - P[3]);
      if (XX >= 0 && YY >= 0) discr = 999;
    } else if (geom == 5) {
      double pi = 4. * atan(1.0);
      double theta = P[5] * pi / 180.;
      double costh = cos(theta), sinth = sin(theta);
      double Rvx = xx - P[1], Rvy = yy - P[2];
      double avx = costh, avy = sinth, bvx = -sinth, bvy = costh;
      double Ra = avx * Rvx + avy * Rvy, Rb = bvx * Rvx + bvy * Rvy;
      if (Ra >= 0.0 && Ra <= P[3] && Rb >= 0.0 && Rb <= P[4]) discr = 999;
    } else if (geom == 101 || geom == 102) {
      int ipx = (int)((xx - d->xmin) / d->dx + 1), ipy = (int)((yy - d->ymin) / d->dy + 1);
      double xmax = d->xmin + (d->nx - 1) * d->dx, ymax = d->ymin + (d->ny - 1) * d->dy;
      if (d->xmax_set) {
        xmax = d->xmax;
        ymax = d->ymax;
      }
      if (ipx == d->nx && fabs(xmax - xx) <= 1.e-3) ipx = ipx - 1;
      if (ipy == d->ny && fabs(ymax - yy) <= 1.e-3) ipy = ipy - 1;
      int64_t npg = (int64_t)d->nx * d->ny, n1 = (int64_t)(ipy - 1) * d->nx + ipx - 1;
      int64_t nn[4] = {n1, n1 + 1, n1 + 1 + d->nx, n1 + d->nx};
      for (int k = 0; k < 4; k++)
        if (nn[k] >= 0 && nn[k] < npg && d->zone_nodes && ((d->zone_nodes[nn[k]] >> z) & 1)) discr = 999;
    }
    if (discr >= 0) return z;
  }
  return -1;
}

void gfo_which_zone(const gfo_dem *d, const double *zc, int n, const double *xy, int btype, int32_t *out) {
  for (int i = 0; i < n; i++)
    out[i] = zc ? which_zone(d, zc, 1, xy[2 * i], xy[2 * i + 1], btype)
                : which_zone(d, d->topol, 9, xy[2 * i], xy[2 * i + 1], btype);
}

void gfo_topo_lookup(const gfo_dem *d, int n, const double *xy, double *out9) {
  for (int i = 0; i < n; i++) {
    topo_pt t;
    get_z_topo(d, xy[2 * i], xy[2 * i + 1], 1, &t);
    double v[9] = {t.Zp, t.Zx, t.Zy, t.t4, t.t5, t.t6, t.t7, t.t8, t.t9};
    memcpy(out9 + 9 * i, v, sizeof(v));
  }
}

/* Fdiff_derivs (2:2812-2877): nodal Zx..Zyy by central differences,      */
/* zero on the boundary and on nodes outside the model.                  */
void gfo_fdiff_derivs(int nx, int ny, double dxg, double dyg, const int32_t *in_model,
                      double *topol /*[npoig][9], Z in col 0 on entry*/) {
  double del2x = 2. * dxg, del2y = 2. * dyg, delx2 = dxg * dxg, dely2 = dyg * dyg,
         delxy = dxg * dyg;
  for (int64_t k = 0; k < (int64_t)nx * ny; k++)
    for (int c = 1; c < 9; c++) topol[k * 9 + c] = 0.0;
  for (int iy = 2; iy <= ny - 1; iy++) {
    for (int ix = 2; ix <= nx - 1; ix++) {
      int64_t iA = (int64_t)(iy - 1) * nx + ix - 1;
      if (in_model && !in_model[iA]) continue;
      int64_t iAL = iA - 1, iAR = iA + 1, iAT = iA + nx, iATL = iAT - 1, iATR = iAT + 1,
              iAB = iA - nx, iABL = iAB - 1, iABR = iAB + 1;
#define Z(n) topol[(n)*9]
      double zx = (Z(iAR) - Z(iAL)) / del2x;
      double zy = (Z(iAT) - Z(iAB)) / del2y;
      double rt = sqrt(1. + zx * zx + zy * zy);
      double zxx = (Z(iAL) - 2. * Z(iA) + Z(iAR)) / delx2;
      double zyy = (Z(iAB) - 2. * Z(iA) + Z(iAT)) / dely2;
      double zxy = (Z(iATR) - Z(iABR) - Z(iATL) + Z(iABL)) / (4. * delxy);
#undef Z
      double *t = topol + iA * 9;
      t[1] = zx;
      t[2] = zy;
      t[3] = 1. + zx * zx;
      t[4] = zx * zy;
      t[5] = 1. + zy * zy;
      t[6] = zxx / rt;
      t[7] = zxy / rt;
      t[8] = zyy / rt;
    }
  }
}

/* ------------------------------------------------------------------ */
/* a4. sum_density_Tools (5a:1763-1859)                                 */
/* ------------------------------------------------------------------ */

static double self_w(int skf, double hsml) {
  double w, a, b;
  gfo_kernel(skf, 0.0, 0.0, 0.0, hsml, &w, &a, &b);
  return w;
}

#define PAIR_K(kk) (p->rev_order ? pv->n - 1 - (kk) : (kk))

static void sum_density(const gfo_params *p, int n, const pairvec *pv, const double *pw,
                        const double *mass, const double *hsml, const int *flag, double *rho) {
  double *wi = NULL;
  if (p->nor_density) {
    wi = (double *)calloc(n, sizeof(double));
    for (int i = 0; i < n; i++)
      if (!flag[i]) wi[i] = self_w(p->skf, hsml[i]) * mass[i] / rho[i];
    for (int64_t kk = 0; kk < pv->n; kk++) {
      int64_t k = PAIR_K(kk);
      int i = pv->i[k], j = pv->j[k];
      wi[i] = wi[i] + mass[j] / rho[j] * pw[k];
      wi[j] = wi[j] + mass[i] / rho[i] * pw[k];
    }
  }
  for (int i = 0; i < n; i++)
    if (!flag[i]) rho[i] = self_w(p->skf, hsml[i]) * mass[i];
  for (int64_t kk = 0; kk < pv->n; kk++) {
    int64_t k = PAIR_K(kk);
    int i = pv->i[k], j = pv->j[k];
    rho[i] = rho[i] + mass[j] * pw[k];
    rho[j] = rho[j] + mass[i] * pw[k];
  }
  if (p->nor_density) {
    for (int i = 0; i < n; i++)
      if (!flag[i]) rho[i] = rho[i] / wi[i];
    free(wi);
  }
}

/* ------------------------------------------------------------------ */
/* a8. basal_friction_Ext_SW (6:9355-9901) + a6 ExtForces_SW(0) body    */
/* (6:8532-8657) for one particle.  Returns ex (force per unit mass)    */
/* and du; writes u back as the reference does at 6:9675.               */
/* ------------------------------------------------------------------ */

typedef struct {
  double ex, ey, du;
  double er; /* erosion_rate of the call (0 without erosion) */
} extout;

/* Get_Erosion_Rate_SW (6:10289-10404), main-region laws, single phase.  h,
   vmod, vxx, vyy, tanfi and frictx/fricty are the values ExtForces_SW's
   shared variables hold after basal_friction_Ext_SW. */
static double erosion_rate(const derived_t *D, double h, double vmod, double vxx, double vyy, double Zpx,
                           double Zpy, double tanfi, double frictx, double fricty) {
  double er = 0.0;
  if (D->eros_law == 1) {
    er = D->keros * h * vmod;
  } else if (D->eros_law == 2) {
    const double Cstar = D->eros_c23, denss = D->denss, densw = D->densw_e;
    double vol_conc = (D->dens - densw) / (denss - densw);
    if (vmod > 1.e-5) {
      double cosa = vxx / vmod, sina = vyy / vmod;
      double gradZv = Zpx * cosa + Zpy * sina;
      double theta = atan2(-gradZv, 1.);
      double xnum = (denss - densw) * vol_conc;
      double xden = xnum + densw;
      xnum = xnum * tanfi;
      double tanthetae = xnum / xden;
      double thetae = atan(tanthetae);
      if (theta > thetae) er = D->keros * Cstar * vmod * tan(theta - thetae);
    }
  } else if (D->eros_law == 3) {
    const double C_basal = D->eros_c23, tanfi_basal = D->eros_c24, beta_basal = D->eros_c25;
    if (vmod > 1.e-5) {
      double dens_bar = D->dens; /* ic_fct_dens = 0 (6:9007) */
      double Taub = dens_bar * sqrt(frictx * frictx + fricty * fricty);
      double Tauf = C_basal + D->dens_bar_eff * D->cgra * h * (1. - beta_basal) * tanfi_basal;
      if (Taub > Tauf && vmod > 1.e-3) er = D->keros * (Taub - Tauf) / D->dens_bar_eff;
    }
  }
  return er;
}

/* Get_Erosion_Rate_SW's erosion zones (6:10407-10503, single phase): in an
   erosion zone its own law replaces er (law 2 with thetae = atan2(xden,
   xnum); law 3 with Taub = dens h |frict|, Tauf = c + dens g h (1 - beta)
   tan(phi_b)); geometry 101 / 102 erodes only where the cell's nodes hold
   available depth (current topol(15)).  zstate: the current [npoig][3]. */
static double zone_erosion_rate(const derived_t *D, const gfo_dem *dem, const double *zstate, double xx, double yy,
                                double er, double h, double vmod, double vxx, double vyy, double Zpx, double Zpy,
                                double tanfi, double frictx, double fricty) {
  const int64_t npg = (int64_t)dem->nx * dem->ny;
  const int z = which_zone(dem, zstate, 3, xx, yy, 2);
  if (z < 0) return er;
  const gfo_zone *Z = &dem->zones[z];
  const double *P = Z->props;
  const int eros_geom = (int)(P[0] + 0.01);
  if (Z->law == 0) {
    er = 0;
  } else if (Z->law == 1) {
    double Keros = P[6];
    er = Keros * h * vmod;
  } else if (Z->law == 2) {
    double Keros = P[6], Cstar = P[7], denss = D->denss, densw = D->densw_e;
    double vol_conc = (D->dens - densw) / (denss - densw);
    er = 0.0;
    if (vmod > 1.e-5) {
      double cosa = vxx / vmod, sina = vyy / vmod;
      double gradZv = Zpx * cosa + Zpy * sina;
      double theta = atan2(-gradZv, 1.);
      double xnum = (denss - densw) * vol_conc;
      double xden = xnum + densw;
      xnum = xnum * tanfi;
      double thetae = atan2(xden, xnum);
      if (theta > thetae) er = Keros * Cstar * vmod * tan(theta - thetae);
    }
  } else if (Z->law == 3) {
    double Keros = P[6], C_basal = P[7], tanfi_basal = P[8], beta_basal = P[9];
    er = 0.0;
    if (vmod > 1.e-5) {
      double Taub = D->dens * h * sqrt(frictx * frictx + fricty * fricty);
      double Tauf = C_basal + D->dens * D->cgra * h * (1. - beta_basal) * tanfi_basal;
      if (Taub > Tauf && vmod > 1.e-3) er = Keros * (Taub - Tauf) / D->dens;
    }
  }
  if (eros_geom == 101 || eros_geom == 102) {
    int ipx = (int)((xx - dem->xmin) / dem->dx + 1), ipy = (int)((yy - dem->ymin) / dem->dy + 1);
    int64_t n1 = (int64_t)(ipy - 1) * dem->nx + ipx - 1;
    int64_t nn[4] = {n1, n1 + 1, n1 + dem->nx, n1 + dem->nx + 1};
    double auxt = 0.0;
    for (int k = 0; k < 4; k++) auxt = auxt + ((nn[k] >= 0 && nn[k] < npg) ? zstate[3 * nn[k] + 2] : 0.0);
    if (auxt < 1.e-6) er = 0.0;
  }
  return er;
}

/* tau0 carries the Bingham Tau0 between particles: ExtForces_SW applies
   cf4b = Tau0/dens after every nfric=4 branch (6:9619-9622) but sets Tau0
   only in the intermediate regime (6:9616), so the viscous and cohesive
   regimes use the Tau0 of the last particle, in particle order, that was in
   the intermediate regime.  The compiled reference behaves this way (checked
   against its dumps); the value carries across calls, starting from 0. */
/* acc0: NULL inside the RK stages (ExtForces_SW(0) clears exdvxdt first,
   6:8482-8485); after them (ExtForces_SW(1), erosion) the particle's
   exdvxdt of the last stage, which the gravity term is added to and the
   limiter reads. */
/* nsoil: (mass - mass0)/mass of the particle, read by nfric = 3 only. */
static void ext_force_one(const gfo_params *p, const derived_t *D, const gfo_dem *dem, double xx,
                          double yy, double vx1, double vy1, double rho_i, double *u_i,
                          double uaux1, double inx, double iny, double *tau0, double nsoil,
                          const double *acc0, extout *o) {
  const double dt = p->dt;
  double h, hu, hv, hinv, vmod, hvmod, hvmodinv;
  double vxx = vx1, vyy = vy1;
  double gravz = D->cgra; /* Get_Grav_Ext_SW sets it per particle when ic_cosTheta = 1 */
  double frictx = 0, fricty = 0, TotFrict = 0, cf5 = 0, hfrict, pwp;
  int nfric = D->nfric, if_frictional = D->if_frictional, if_V2byR = D->if_V2byR;
  double xi_voellmy = D->xi_voellmy;
  double exx = acc0 ? acc0[0] : 0.0, exy = acc0 ? acc0[1] : 0.0, du = 0.0;
  double tanfi_m = 0.0; /* ExtForces_SW's tanfi after basal_friction_Ext_SW */

  h = rho_i;
  hu = h * vx1;
  hv = h * vy1;
  if (h < D->comp) { /* 6:8578-8596 */
    h = D->comp;
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

  /* Get_Grav_Ext_SW (6:9257-9299): g along the terrain normal when
     ic_cosTheta = 1, else vertical g and the slope terms */
  topo_pt tp;
  get_z_topo(dem, xx, yy, if_frictional, &tp);
  double gravx, gravy;
  if (D->ic_cosTheta == 1) {
    const double deng = sqrt(1. + tp.Zx * tp.Zx + tp.Zy * tp.Zy);
    gravz = D->cgra / deng;
    gravx = -D->cgra * tp.Zx / deng;
    gravy = -D->cgra * tp.Zy / deng;
  } else {
    gravx = -D->cgra * tp.Zx;
    gravy = -D->cgra * tp.Zy;
  }
  exx = exx + gravx;
  exy = exy + gravy;

  /* basal_friction_Ext_SW */
  if (p->icpwp == 1) pwp = *u_i;
  else if (p->icpwp == 2 || p->icpwp == 3) pwp = uaux1;
  else pwp = 0.0;
  hfrict = fmax(h, D->hfrict0);

  if (nfric == 1) { /* 6:9426 */
    double cf1 = D->cgra * xi_voellmy * xi_voellmy; /* cmanning = const(3) */
    double hp = pow(hfrict, D->c73);
    frictx = cf1 * hu * hvmod / hp;
    fricty = cf1 * hv * hvmod / hp;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 2) { /* 6:9432 */
    double cf2a = D->constK / (8. * D->dens);
    frictx = cf2a * hu / (hfrict * hfrict);
    fricty = cf2a * hv / (hfrict * hfrict);
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 3) { /* Bingham erosion 6:9443-9488 */
    double frictxx, frictyy;
    const double nsoilB = nsoil; /* (mass - mass0)/mass, 6:9451 */
    const double csoilB = p->cst[3];
    const double fsoilB = 1. - exp(-csoilB * nsoilB);
    const double constKB = D->constK * fsoilB, tauy0B = D->tauy0 * fsoilB;
    const double densB = 1000. + (D->dens - 1000.) * nsoilB;
    frictx = 0.0;
    fricty = 0.0;
    if (nsoilB > 0.01) {
      double a_hat = (6. * constKB * hvmod) / (tauy0B * (hfrict * hfrict));
      if (a_hat > 50) { /* viscous */
        double cf4b = 3 * constKB / densB;
        frictx = cf4b * hu / (hfrict * hfrict);
        fricty = cf4b * hv / (hfrict * hfrict);
      } else if (a_hat < 0.01) { /* cohesive */
        double cf4b = tauy0B / densB;
        frictx = cf4b * hu * hvmodinv;
        fricty = cf4b * hv * hvmodinv;
      } else { /* intermediate */
        double a2 = 48., b2 = -32. * a_hat - 114., c2 = 65.;
        double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
        double cf4b = (tauy0B / xh) / densB;
        frictx = cf4b * hu * hvmodinv;
        fricty = cf4b * hv * hvmodinv;
      }
    }
    const double cmanning = p->cst[8];
    const double cf1 = D->cgra * cmanning * cmanning;
    const double hp = pow(hfrict, D->c73);
    frictxx = cf1 * hu * hvmod / hp;
    frictyy = cf1 * hv * hvmod / hp;
    if (fabs(frictxx) > fabs(frictx)) frictx = frictxx;
    if (fabs(frictyy) > fabs(fricty)) fricty = frictyy;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 23) { /* Herschel-Bulkley, exact root (6:9538-9545, 6:9576-9605, 6:9619-9622) */
    const double Tauy0 = D->tauy0, constK = D->constK, m_vexp = D->visco;
    const double Cm = m_vexp / ((m_vexp + 1) * (2 * m_vexp + 1));
    double a_hat = (1. / Cm) * pow(constK / Tauy0, 1. / m_vexp) * (hvmod / (hfrict * hfrict));
    double a2, b2, c2;
    if (m_vexp == 1) {
      a2 = 48.;
      b2 = -32. * a_hat - 114.;
      c2 = 65.;
    } else if (m_vexp == 2) {
      a2 = 10.;
      b2 = -a_hat - 19.2188;
      c2 = 9.1875;
    } else { /* m = 1/3 or 1/2 (other exponents stop the reference) */
      a_hat = (4.5 * constK * hvmod) / (Tauy0 * (hfrict * hfrict));
      a2 = 48.;
      b2 = -32. * a_hat - 114.;
      c2 = 65.;
    }
    double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
    double Tau0 = Tauy0 / xh;
    double cf4b = Tau0 / D->dens;
    frictx = cf4b * hu * hvmodinv;
    fricty = cf4b * hv * hvmodinv;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 4 || nfric == 24 || nfric == 25) { /* 6:9538-9623 */
    double Tauy0 = D->tauy0, m_vexp = D->visco, Tau0 = *tau0;
    double constMu_CF = D->constK;
    if (nfric == 4) {
      double constK = D->constK;
      double a_hat = (6. * constK * hvmod) / (Tauy0 * (hfrict * hfrict));
      if (!(a_hat > 50) && !(a_hat < 0.01)) { /* intermediate regime */
        double a2 = 48., b2 = -32. * a_hat - 114., c2 = 65.;
        double xh = (-b2 - sqrt(b2 * b2 - 4. * a2 * c2)) / (2. * a2);
        Tau0 = Tauy0 / xh;
        *tau0 = Tau0;
      }
      double cf4b = Tau0 / D->dens; /* 6:9619-9622, every regime */
      frictx = cf4b * hu * hvmodinv;
      fricty = cf4b * hv * hvmodinv;
    } else {
      Tau0 = constMu_CF * pow(vmod / hfrict, m_vexp);
      double Cm = (nfric == 24) ? 1. : pow((2 * m_vexp + 1) / m_vexp, m_vexp);
      Tau0 = Tauy0 + Tau0 * Cm;
      double cf4b = Tau0 / D->dens;
      frictx = cf4b * hu * hvmodinv;
      fricty = cf4b * hv * hvmodinv;
    }
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (if_frictional == 1) { /* 6:9624-9789 */
    double cv2byR = 0.0, tanfi;
    if (if_V2byR == 1) {
      vxx = hu * hinv;
      vyy = hv * hinv;
      vmod = sqrt(vxx * vxx + vyy * vyy);
      cv2byR = 0.0;
      if (vmod >= 1e-5) {
        double cosa = vxx / vmod, sina = vyy / vmod;
        double cos2 = cosa * cosa, sin2 = sina * sina, cbs = sina * cosa;
        double xnum = tp.t7 * cos2 + tp.t8 * cbs + tp.t9 * sin2;
        double xden = tp.t4 * cos2 + tp.t5 * cbs + tp.t6 * sin2;
        cv2byR = vmod * vmod * (xnum / xden);
      }
    }
    tanfi = D->tanfi8;
    if (nfric == 8 || nfric == 26 || nfric == 27) {
      xi_voellmy = 0.0;
      if (nfric == 8 || nfric == 26) {
        double b_muI = p->cst[3], r_muI = p->cst[6], n_muI = p->cst[7];
        double Froude = vmod / sqrt(gravz * h);
        double ff = (1 + b_muI * pow(Froude * r_muI / h, n_muI));
        tanfi = tanfi * ff;
      } else {
        double Mu2 = p->cst[3], r_muI = p->cst[6], Inert0 = p->cst[7];
        double Froude = vmod / sqrt(gravz * h);
        double Inert = Froude * r_muI / h;
        tanfi = tanfi + (Mu2 - tanfi) / (1.0 + (Inert0 / Inert));
      }
    }
    double hhh = rho_i;
    if (pwp > D->fpwp_max * hhh) pwp = D->fpwp_max * hhh;
    if (pwp < D->fpwp_min * hhh) pwp = D->fpwp_min * hhh;
    *u_i = pwp; /* 6:9675 */
    if (p->icpwp == 1 || p->icpwp == 2 || p->icpwp == 3) {
      double factpw = 1. - pwp * hinv;
      if (factpw < D->fact_pwp_min) factpw = D->fact_pwp_min;
      if (factpw > D->fact_pwp_max) factpw = D->fact_pwp_max;
      tanfi = tanfi * factpw;
    }
    /* friction zones of the terrain: Get_Topo_Basal_Friction_SW (6:9697-9702,
       6:10515-10545) -- Tandelta = basal_props(7) of the zone, if >= 0 */
    if (dem->nzones) {
      const int z = D->eros ? which_zone(dem, D->eros, 3, xx, yy, 1) : which_zone(dem, dem->topol, 9, xx, yy, 1);
      if (z >= 0 && dem->zones[z].props[6] >= 0) tanfi = dem->zones[z].props[6];
    }
    tanfi_m = tanfi;
    double gtanfi = (gravz + cv2byR) * tanfi * D->fact_dens;
    if (gtanfi < 0.0) gtanfi = 0.0;
    cf5 = gtanfi;
    frictx = cf5 * hu * hvmodinv * h;
    fricty = cf5 * hv * hvmodinv * h;
    if (xi_voellmy > 0.0) { /* Voellmy 6:9719-9749 */
      vxx = vx1;
      vyy = vy1;
      vmod = vxx * vxx + vyy * vyy;
      vmod = (vmod > 0.0) ? sqrt(vmod) : 0.0;
      frictx = frictx + D->cgra * vmod * vxx / xi_voellmy;
      fricty = fricty + D->cgra * vmod * vyy / xi_voellmy;
    } else if (xi_voellmy < 0.) { /* 6:9751 */
      vxx = vx1;
      vyy = vy1;
      vmod = vxx * vxx + vyy * vyy;
      vmod = (vmod > 0.0) ? sqrt(vmod) : 0.0;
      frictx = frictx + D->constMu_CF * vmod * vxx / (hfrict * hfrict);
      fricty = fricty + D->constMu_CF * vmod * vyy / (hfrict * hfrict);
    }
    if (nfric == 9 || nfric == 10) { /* Perzyna 6:9767-9787 */
      double C6 = p->cst[6], C7 = p->cst[7], C8 = p->cst[8], C67 = C6 + C7;
      double frict_abs = sqrt(frictx * frictx + fricty * fricty);
      double fact9 = pow(2. * C8 * vmod / hfrict, 1. / C67) * pow(frict_abs, -C6 / C67);
      if (p->cst[47] > 1.e-6) { /* 6:9780-9784 */
        double alpha_slip = p->cst[47];
        fact9 = fact9 * ((1 - alpha_slip) / (1 + alpha_slip));
      }
      frictx = frictx * (1 + fact9);
      fricty = fricty * (1 + fact9);
    }
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 11) { /* Bagnold 6:9791 */
    vxx = hu * hinv;
    vyy = hv * hinv;
    frictx = D->K_Bagnold * vxx * vmod / (hfrict * hfrict);
    fricty = D->K_Bagnold * vyy * vmod / (hfrict * hfrict);
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  } else if (nfric == 22) { /* simplified Bingham + power law 6:9803-9816 */
    /* vxx, vyy are the particle's velocity and vmod = |hv|/h (6:8568-8589);
       a negative component adds nothing and the y term goes to frictx, as
       the reference writes it */
    double tb_abs = D->tauy0 + (D->constK * pow(vmod / hfrict, D->visco));
    tb_abs = tb_abs / D->dens;
    if (vxx >= 1.e-4) frictx = tb_abs * vxx / vmod;
    if (vyy >= 1.e-4) frictx = frictx + tb_abs * vyy / vmod;
    TotFrict = sqrt(frictx * frictx + fricty * fricty);
  }

  /* limiter (6:9836-9895) */
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
      if (if_frictional == 1) ctf = cf5;
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

  if (p->icpwp == 1) { /* 6:8648-8652 */
    double betaw = D->Bfact * hinv * hinv / D->hrelpw2;
    du = -(*u_i / dt) * (1 - exp(-betaw * dt));
  }
  exx = exx - frictx / h;
  exy = exy - fricty / h;
  o->er = 0.0;
  if (D->eros_on) { /* 6:8664-8693; the momentum term only inside the stages */
    o->er = erosion_rate(D, h, vmod, vxx, vyy, tp.Zx, tp.Zy, tanfi_m, frictx, fricty);
    if (dem->nzones)
      o->er = zone_erosion_rate(D, dem, D->eros, xx, yy, o->er, h, vmod, vxx, vyy, tp.Zx, tp.Zy, tanfi_m, frictx,
                                fricty);
    if (!acc0) {
      exx = exx - o->er * vx1 * hinv * D->slip_fact;
      exy = exy - o->er * vy1 * hinv * D->slip_fact;
    }
  }
  o->ex = exx;
  o->ey = exy;
  o->du = du;
}

/* ------------------------------------------------------------------ */
/* a13. Check_Out_Domain (7:1764-1788): sticky flag                    */
/* ------------------------------------------------------------------ */
static inline int out_of_domain(const gfo_dem *d, double x, double y) {
  double xmax = d->xmax_set ? d->xmax : d->xmin + (d->nx - 1) * d->dx;
  double ymax = d->xmax_set ? d->ymax : d->ymin + (d->ny - 1) * d->dy;
  return ((x - d->xmin) * (x - xmax) > 0.0) || ((y - d->ymin) * (y - ymax) > 0.0);
}

/* ------------------------------------------------------------------ */
/* Input_SW initialisation for icpwp != 0 (6:1428-1452, 6:1705-1852):  */
/* re-sum h from the initial positions, then initialise the pwp.        */
/* ------------------------------------------------------------------ */
int gfo_init(const gfo_params *p, const gfo_dem *dem, int n, const double *x, double *h,
             const double *mass, const double *hsml, double *u, double *uaux, int32_t *flag) {
  for (int i = 0; i < n; i++)
    if (out_of_domain(dem, x[2 * i], x[2 * i + 1])) flag[i] = 1;
  if (p->icpwp != 0 && p->summation_density) {
    pairvec pv = {0, 0, NULL, NULL};
    build_pairs(n, p->skf, x, hsml, flag, 0, &pv);
    double *pw = (double *)malloc(sizeof(double) * (pv.n ? pv.n : 1));
    int scale_k = (p->skf == 1 || p->skf == 4) ? 2 : 3;
    for (int64_t k = 0; k < pv.n; k++) {
      int i = pv.i[k], j = pv.j[k];
      double dx, dy, r, mh, gx, gy;
      pair_test(scale_k, x[2 * i], x[2 * i + 1], x[2 * j], x[2 * j + 1], hsml[i], hsml[j], &dx,
                &dy, &r, &mh);
      gfo_kernel(p->skf, r, dx, dy, mh, &pw[k], &gx, &gy);
    }
    sum_density(p, n, &pv, pw, mass, hsml, flag, h);
    free(pw);
    free(pv.i);
    free(pv.j);
  }
  double alfs = p->cst[14];
  if (p->icpwp == 2 && p->nuaux > 1) { /* Initialize_Pwp, icpwp = 2 (6:1769-1803) */
    const int N = p->nuaux;
    const double pi = 4. * atan(1.0);
    for (int i = 0; i < n; i++) {
      double hh = h[i], *U = uaux + (int64_t)i * N;
      for (int k = 0; k < N; k++) {
        double xi, v = 0.0;
        switch (p->pwp0_type) {
          case 0: v = p->pwp0rel * hh; break;
          case 1: xi = 1. * k / (N - 1); v = (1 - xi) * p->pwp0rel * hh; break;
          case 2: xi = (k + 1 - 1.) / (N - 1.); v = cos(pi * xi / 2) * p->pwp0rel * hh; break;
          case 3: v = (k == 0) ? p->pwp0rel * hh : 0.0; break;
          case 4: xi = (k + 1 - 1.) / (N - 1.); v = (xi < p->xirel) ? hh * p->pwp0rel * (1 - xi) : 0.0; break;
          default: return -2;
        }
        U[k] = v;
      }
      u[i] = U[0];
    }
    return 0;
  }
  if (p->icpwp == 1) {
    for (int i = 0; i < n; i++) u[i] = p->pwp0rel * (h[i] * alfs);
  } else if (p->icpwp == 3 || p->icpwp == 2) {
    for (int i = 0; i < n; i++) {
      uaux[i] = p->pwp0rel * (h[i] * alfs);
      u[i] = uaux[i];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* a10. RK45 (7:746-1005) and the RK2 (Heun) extension                  */
/* ------------------------------------------------------------------ */
/* Adaptive_dt_sph_SW (6:6521-6664) for ic_adapt_dt_sph = 1, 2, 4, 10: the
   smallest per-particle bound over the active particles (clamping their depth
   to h_inf_SW first where the mode reads it), then dt_sph = dtmin if that is
   smaller, else min(1.1 dt_sph, dtmin). */
static double adaptive_dt(const gfo_params *p, const gfo_dem *dem, int n, const double *x, const double *v,
                          double *rho, const double *hsml, const int32_t *flag, double dt) {
  const int mode = p->adapt_dt;
  const double cgra = p->cst[1];
  const double dxg_mod = sqrt(dem->dx * dem->dx + dem->dy * dem->dy); /* **0.5: a square root */
  double dtmin = 1.e10;
  if (mode == 7) { /* one-dimensional free surface, particles in index order (6:6597-6607) */
    for (int i = 0; i + 1 < n; i++) {
      double Le = fabs(x[2 * (i + 1)] - x[2 * i]);
      double vv = fabs(v[2 * (i + 1)] + v[2 * i]) / 2.;
      if (vv > 1.e-6) {
        double cc = 0.5 * Le / vv;
        dtmin = fmin(cc, dtmin);
      }
    }
  }
  for (int i = 0; i < n && mode != 7; i++) {
    if (flag[i]) continue;
    double vv = sqrt(v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1]), cc;
    if (mode == 1 || mode == 2 || mode == 10) {
      if (rho[i] < p->h_inf) rho[i] = p->h_inf;
      const double c = sqrt(cgra * rho[i]);
      cc = fmax(1.e-6, c);
      if (mode == 2) cc = cc + vv;
      cc = hsml[i] / (3 * cc);
      dtmin = fmin(cc, dtmin);
    }
    if (mode == 4 || mode == 10) {
      cc = dxg_mod / vv;
      dtmin = fmin(cc, dtmin);
    }
  }
  if (dtmin < dt) return dtmin;
  return fmin(1.1 * dt, dtmin);
}

double gfo_adaptive_dt(const gfo_params *p, const gfo_dem *dem, int n, const double *x, const double *v,
                       double *rho, const double *hsml, const int32_t *flag, double dt) {
  return adaptive_dt(p, dem, n, x, v, rho, hsml, flag, dt);
}

/* Get_Pwp_SW for npwp > 1 (6:10966-11174), single phase without topo
   boundary zones: per particle an explicit FTCS consolidation of the column
   Uaux(1:N) over the saturated depth h*alphaS, in ntpwp sub-steps, with the
   loading term dh/ntpwp*(1-z/h), the clamps z-dependent, and the bottom /
   top conditions (1 value, 2 gradient).  The sub-step count is capped at
   PWP_MAX_SUBSTEPS (the reference has no cap; a cap is only reached for a
   vanishing depth). */
#define PWP_MAX_SUBSTEPS (1 << 16)
static void pwp_column(const gfo_params *p, const derived_t *D, const gfo_dem *dem, const double *x, int n, int N,
                       double dt, const double *rho, const double *drho, const int32_t *flag, double *uaux) {
  const double Cv = D->Cv, alphaS = p->cst[14];
  double *pwp0 = (double *)malloc(sizeof(double) * N);
  for (int i = 0; i < n; i++) {
    double *pwp = uaux + (int64_t)i * N;
    if (flag[i]) {
      for (int k = 0; k < N; k++) pwp[k] = 0.0;
      continue;
    }
    int LBc1 = p->pwp_bc_type[0];
    const int LBc2 = p->pwp_bc_type[1];
    double VBc1 = p->pwp_bc_val[0];
    if (dem->nzones) { /* pore-pressure zones of the terrain (basal type 3, 6:10987-11001) */
      const int z = D->eros ? which_zone(dem, D->eros, 3, x[2 * i], x[2 * i + 1], 3)
                            : which_zone(dem, dem->topol, 9, x[2 * i], x[2 * i + 1], 3);
      if (z >= 0) {
        LBc1 = (int)dem->zones[z].props[6];
        VBc1 = dem->zones[z].props[7];
      }
    }
    const double VBc2 = p->pwp_bc_val[1];
    double fpwp_min = D->fpwp_min;
    const double fpwp_max = D->fpwp_max;
    const double hh = rho[i], hpwp = hh * alphaS, dhh = drho[i], dpwp_dn = 0.0;
    if (VBc1 < -900) { /* 6:11053-11060 */
      VBc1 = -((p->cst[18] < 1.e-3 ? 1000.0 : p->cst[18]) / D->dens_bar_eff);
      fpwp_min = VBc1;
    }
    const double dxpwp = hpwp / (N - 1), dxpwp2 = dxpwp * dxpwp;
    double dtpwp = (Cv > 1.e-20) ? dxpwp2 / (2 * Cv) : dxpwp2 / (2.e-20);
    dtpwp = 1.0 * dtpwp; /* safetpwp = 1 */
    double q = ceil(dt / dtpwp);
    const int ntpwp = (q >= 1.0 && q <= PWP_MAX_SUBSTEPS) ? (int)q : (q >= 1.0 ? PWP_MAX_SUBSTEPS : 1);
    dtpwp = dt / (double)ntpwp;
    const double dhhstep = dhh / ntpwp, dpwp_dn_step = dpwp_dn / ntpwp;
    for (int it = 0; it < ntpwp; it++) {
      for (int k = 0; k < N; k++) pwp0[k] = pwp[k];
      for (int k = 1; k < N - 1; k++) { /* ip = k + 1 */
        const double zz = dxpwp * k, zzhh = hh - zz, xi = zz / hh;
        const double sumdhh = dhhstep * (1. - xi);
        pwp[k] = pwp0[k] + (Cv * dtpwp / dxpwp2) * (pwp0[k - 1] - 2. * pwp0[k] + pwp0[k + 1]) + sumdhh +
                 dpwp_dn_step;
        if (pwp[k] > (zzhh * fpwp_max)) pwp[k] = (zzhh * fpwp_max);
        else if (pwp[k] < (zzhh * fpwp_min)) pwp[k] = (zzhh * fpwp_min);
      }
      if (LBc1 == 1) pwp[0] = VBc1 * hh * alphaS;
      else if (LBc1 == 2) pwp[0] = (4. * pwp[1] - pwp[2] - 2. * VBc1 * dxpwp) / 3.;
      if (pwp[0] > hh) pwp[0] = hh;
      if (LBc2 == 1) pwp[N - 1] = VBc2 * hh;
      else if (LBc2 == 2) pwp[N - 1] = (4. * pwp[N - 2] - pwp[N - 3] + 2. * VBc2 * dxpwp) / 3.;
      if (pwp[N - 1] > 0.0) pwp[N - 1] = 0.0;
    }
  }
  free(pwp0);
}

/* a5. IntForces_SW (6:6669-6941) with the frozen pair geometry of the step:
   the pressures p(i) of the active particles -- 0.5 g h^2, with g along
   the terrain normal at the particle when ic_cosTheta = 1 (6:6774-6783),
   times K0_active / K0_pasive by the sign of the velocity divergence when
   ic_K0 = 1 (6:6736-6760, 6:6857-6863) -- then the pair sum.  p keeps the
   values of particles that are out of the domain (not recomputed).  Called
   once per step when neither applies (the force is then the same in every
   stage), else in every stage with that stage's positions and velocities. */
static void pressure_force(const gfo_params *p, const derived_t *D, const gfo_dem *dem, int n, const pairvec *pv,
                           const double *pdw, const double *mass, const double *rhoF, const int32_t *flag,
                           const double *x, const double *v, double *p_, double *divv, double *inf) {
  if (p->ic_k0) { /* divvK0 over the pairs, stage velocities (6:6736-6756) */
    memset(divv, 0, sizeof(double) * n);
    for (int64_t k = 0; k < pv->n; k++) {
      int i = pv->i[k], j = pv->j[k];
      double vcc = (v[2 * i] - v[2 * j]) * pdw[2 * k];
      vcc = vcc + (v[2 * i + 1] - v[2 * j + 1]) * pdw[2 * k + 1];
      divv[i] = divv[i] + mass[j] * vcc;
      divv[j] = divv[j] + mass[i] * vcc;
    }
  }
  for (int i = 0; i < n; i++) {
    if (flag[i]) continue;
    double cgra05 = D->cgra * 0.5;
    if (D->ic_cosTheta == 1) {
      topo_pt tp;
      get_z_topo(dem, x[2 * i], x[2 * i + 1], 0, &tp);
      const double deng = sqrt(1. + tp.Zx * tp.Zx + tp.Zy * tp.Zy);
      cgra05 = 0.5 * (D->cgra / deng);
    }
    p_[i] = cgra05 * rhoF[i] * rhoF[i];
    if (p->ic_k0) p_[i] = p_[i] * (divv[i] >= 0 ? p->k0_active : p->k0_pasive);
  }
  memset(inf, 0, sizeof(double) * 2 * n);
  for (int64_t kk = 0; kk < pv->n; kk++) {
    int64_t k = p->rev_order ? pv->n - 1 - kk : kk;
    int i = pv->i[k], j = pv->j[k];
    double mI = mass[i], mJ = mass[j];
    for (int d = 0; d < 2; d++) {
      double hh = 0, gw = pdw[2 * k + d];
      if (p->pa_sph == 1) {
        double rhoij = 1.e0 / (rhoF[i] * rhoF[j]);
        hh = -(p_[i] + p_[j]) * gw;
        hh = hh * rhoij;
        inf[2 * i + d] += mJ * hh;
        inf[2 * j + d] -= mI * hh;
      } else if (p->pa_sph == 2) {
        hh = -(p_[i] / (rhoF[i] * rhoF[i]) + p_[j] / (rhoF[j] * rhoF[j])) * gw;
        inf[2 * i + d] += mJ * hh;
        inf[2 * j + d] -= mI * hh;
      } else if (p->pa_sph == 4) {
        inf[2 * i + d] -= ((p_[j] - p_[i]) * gw * mJ / rhoF[j]);
        inf[2 * j + d] -= ((p_[j] - p_[i]) * gw * mI / rhoF[i]);
      }
    }
  }
}

/* Node coordinates coorg(:, ipoig) of the terrain grid (2:285-296). */
/* (mass - mass0)/mass for the Bingham-erosion law (6:9451); 0 for others */
static double nsoil_of(const gfo_params *p, int nfric, const double *mass, int i) {
  if (nfric != 3) return 0.0;
  return (mass[i] - p->mass0[i]) / mass[i];
}

static void node_xy(const gfo_dem *d, int64_t ip, double *xg, double *yg) {
  if (d->coorg) {
    *xg = d->coorg[2 * ip];
    *yg = d->coorg[2 * ip + 1];
  } else {
    *xg = d->xmin + (ip % d->nx) * d->dx;
    *yg = d->ymin + (ip / d->nx) * d->dy;
  }
}

int gfo_run(const gfo_params *p_in, const gfo_dem *dem, int n, double *x, double *v, double *rho,
            double *mass, double *hsml, double *u, double *uaux, int32_t *flag,
            int nsteps, double *tau0_io, double *dt_time_io, double *eros) {
  derived_t D;
  derive(p_in, &D);
  for (int z = 0; z < dem->nzones; z++)
    if (dem->zones[z].type == 2) D.eros_on = 1; /* ic_basal_erosion (6:1288-1290) */
  D.eros = D.eros_on ? eros : NULL;
  if (D.eros_on && !eros) return -2;
  if (D.nfric == 3 && !p_in->mass0) return -3;
  const int64_t npg = (int64_t)dem->nx * dem->ny;
  double *auxv = D.eros_on ? (double *)calloc(2 * npg, sizeof(double)) : NULL;
  double dt = dt_time_io ? dt_time_io[0] : p_in->dt;
  double time_sph = dt_time_io ? dt_time_io[1] : 0.0;
  gfo_params p_step = *p_in;
  const gfo_params *p = &p_step; /* dt_sph of the current step */
  p_step.dt = dt;
  const int nst = (p->integrator == 2) ? 2 : 4;
  const double f1rk4[4] = {0., 0.5, 0.5, 1.0}, f2rk4[4] = {1., 2., 2., 1.0};
  const double f1rk2[2] = {0., 1.0}, f2rk2[2] = {1., 1.};
  const double *f1 = (nst == 4) ? f1rk4 : f1rk2, *f2 = (nst == 4) ? f2rk4 : f2rk2;
  double fin = (nst == 4) ? dt / 6. : dt / 2.;
  const int nuaux = (p->icpwp == 2 || p->icpwp == 3) ? 1 : 0;
  const int ncol = (p->icpwp == 2 && p->nuaux > 1) ? p->nuaux : 1; /* Uaux(1:ncol, ipoin) */
  int scale_k = D.scale_k;

  double *x0 = (double *)malloc(sizeof(double) * 2 * n), *v0 = (double *)malloc(sizeof(double) * 2 * n);
  double *u0 = (double *)malloc(sizeof(double) * n), *rho0 = (double *)malloc(sizeof(double) * n);
  double *S = (double *)malloc(sizeof(double) * n), *inf = (double *)malloc(sizeof(double) * 2 * n);
  double *ex = (double *)malloc(sizeof(double) * 2 * n), *du = (double *)malloc(sizeof(double) * n);
  double *Dx = (double *)malloc(sizeof(double) * 2 * n), *Dv = (double *)malloc(sizeof(double) * 2 * n);
  double *Du = (double *)malloc(sizeof(double) * n), *p_ = (double *)malloc(sizeof(double) * n);
  double *av = (double *)malloc(sizeof(double) * 2 * n);
  double *divv = (double *)calloc(n, sizeof(double));
  pairvec pv = {0, 0, NULL, NULL};
  double *pw = NULL, *pdw = NULL;
  int64_t pcap = 0;
  double tau0_carry = tau0_io ? *tau0_io : 0.0;

  for (int step = 0; step < nsteps; step++) {
    /* Driver_Flow: adaptive_dt_sph before time_integration_sph (8:1610-1616) */
    if (p->adapt_dt) {
      dt = adaptive_dt(p, dem, n, x, v, rho, hsml, flag, dt);
      fin = (nst == 4) ? dt / 6. : dt / 2.;
      p_step.dt = dt;
    }
    /* time_integration_sph: Check_Out_Domain before RK45 (7:232) */
    for (int i = 0; i < n; i++)
      if (out_of_domain(dem, x[2 * i], x[2 * i + 1])) flag[i] = 1;

    /* 1. neighbour list + frozen kernel values at x^n (7:784-799) */
    build_pairs(n, p->skf, x, hsml, flag, 0, &pv);
    if (pv.n > pcap) {
      pcap = pv.n;
      pw = (double *)realloc(pw, sizeof(double) * pcap);
      pdw = (double *)realloc(pdw, sizeof(double) * 2 * pcap);
    }
    for (int64_t k = 0; k < pv.n; k++) {
      int i = pv.i[k], j = pv.j[k];
      double dx, dy, r, mh;
      pair_test(scale_k, x[2 * i], x[2 * i + 1], x[2 * j], x[2 * j + 1], hsml[i], hsml[j], &dx,
                &dy, &r, &mh);
      gfo_kernel(p->skf, r, dx, dy, mh, &pw[k], &pdw[2 * k], &pdw[2 * k + 1]);
    }

    memcpy(x0, x, sizeof(double) * 2 * n);
    memcpy(v0, v, sizeof(double) * 2 * n);
    memcpy(u0, u, sizeof(double) * n);
    memcpy(rho0, rho, sizeof(double) * n);
    memset(Dx, 0, sizeof(double) * 2 * n);
    memset(Dv, 0, sizeof(double) * 2 * n);
    memset(Du, 0, sizeof(double) * n);
    memset(ex, 0, sizeof(double) * 2 * n);
    memset(inf, 0, sizeof(double) * 2 * n);
    memset(du, 0, sizeof(double) * n);

    /* 2. density with frozen w: identical in every stage and after them
          (7:844-848, 7:926-931) */
    memcpy(S, rho, sizeof(double) * n);
    sum_density(p, n, &pv, pw, mass, hsml, flag, S);

    /* 3. pressure force with frozen dwdx (IntForces_SW, 6:6769-6941):
          the same in every stage unless ic_K0 = 1 or ic_cosTheta = 1 */
    double *rhoF = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) rhoF[i] = (S[i] < p->h_inf) ? p->h_inf : S[i];
    const int staged = p->ic_k0 || D.ic_cosTheta == 1;
    if (!staged) pressure_force(p, &D, dem, n, &pv, pdw, mass, rhoF, flag, x, v, p_, divv, inf);

    /* 4. RK stages (7:808-906) */
    for (int s = 0; s < nst; s++) {
      for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        x[2 * i] = x0[2 * i] + f1[s] * dt * v[2 * i];
        x[2 * i + 1] = x0[2 * i + 1] + f1[s] * dt * v[2 * i + 1];
        v[2 * i] = v0[2 * i] + f1[s] * dt * (ex[2 * i] + inf[2 * i]);
        v[2 * i + 1] = v0[2 * i + 1] + f1[s] * dt * (ex[2 * i + 1] + inf[2 * i + 1]);
        u[i] = u0[i] + f1[s] * dt * du[i];
        if (nuaux == 0) {
          double rr = (s == 0) ? rho[i] : rhoF[i];
          if (u[i] > rr) u[i] = rr;
          else if (u[i] < 0.0) u[i] = 0.0;
        }
        if (out_of_domain(dem, x[2 * i], x[2 * i + 1])) flag[i] = 1;
      }
      for (int i = 0; i < n; i++) {
        rho[i] = flag[i] ? rho[i] : S[i];
        if (rho[i] < p->h_inf) rho[i] = p->h_inf; /* 6:6769 applies to all */
      }
      if (staged) pressure_force(p, &D, dem, n, &pv, pdw, mass, rhoF, flag, x, v, p_, divv, inf);
      for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        extout o;
        ext_force_one(p, &D, dem, x[2 * i], x[2 * i + 1], v[2 * i], v[2 * i + 1], rho[i], &u[i],
                      nuaux ? uaux[(int64_t)i * ncol] : 0.0, inf[2 * i], inf[2 * i + 1], &tau0_carry,
                      nsoil_of(p, D.nfric, mass, i), NULL, &o);
        ex[2 * i] = o.ex;
        ex[2 * i + 1] = o.ey;
        du[i] = o.du;
        Dv[2 * i] += f2[s] * (ex[2 * i] + inf[2 * i]);
        Dv[2 * i + 1] += f2[s] * (ex[2 * i + 1] + inf[2 * i + 1]);
        Dx[2 * i] += f2[s] * v[2 * i];
        Dx[2 * i + 1] += f2[s] * v[2 * i + 1];
        Du[i] += f2[s] * du[i];
      }
    }
    /* 5. final update (7:908-922) */
    for (int i = 0; i < n; i++) {
      if (flag[i]) continue;
      x[2 * i] = x0[2 * i] + fin * Dx[2 * i];
      x[2 * i + 1] = x0[2 * i + 1] + fin * Dx[2 * i + 1];
      v[2 * i] = v0[2 * i] + fin * Dv[2 * i];
      v[2 * i + 1] = v0[2 * i + 1] + fin * Dv[2 * i + 1];
      u[i] = u0[i] + fin * Du[i];
    }
    /* 6. final density: rho = S for active particles (7:926-931) */
    double *Drho = rhoF; /* reuse storage */
    for (int i = 0; i < n; i++) {
      if (!flag[i]) rho[i] = S[i];
      Drho[i] = rho[i] - rho0[i];
    }
    for (int i = 0; i < n; i++)
      if (out_of_domain(dem, x[2 * i], x[2 * i + 1])) flag[i] = 1;
    /* 7. ExtForces_SW(1) (7:941): erosion once the step's velocities are
          known (6:8724-8761), then the terrain (6:8790-8806) */
    if (D.eros_on) {
      const double dxg = dem->dx, dyg = dem->dy;
      memset(auxv, 0, sizeof(double) * 2 * npg);
      for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        extout o;
        ext_force_one(p, &D, dem, x[2 * i], x[2 * i + 1], v[2 * i], v[2 * i + 1], rho[i], &u[i],
                      nuaux ? uaux[(int64_t)i * ncol] : 0.0, inf[2 * i], inf[2 * i + 1], &tau0_carry,
                      nsoil_of(p, D.nfric, mass, i), &ex[2 * i], &o);
        const double eb = dt * o.er;
        const double xx = x[2 * i], yy = x[2 * i + 1];
        int ipx = (int)((xx - dem->xmin) / dxg + 1), ipy = (int)((yy - dem->ymin) / dyg + 1);
        int inx0 = ipx - 1 > 1 ? ipx - 1 : 1, iny0 = ipy - 1 > 1 ? ipy - 1 : 1;
        int inx1 = ipx + 2 < dem->nx ? ipx + 2 : dem->nx, iny1 = ipy + 2 < dem->ny ? ipy + 2 : dem->ny;
        for (int inx = inx0; inx <= inx1; inx++)
          for (int iny = iny0; iny <= iny1; iny++) {
            int64_t ip = (int64_t)(iny - 1) * dem->nx + (inx - 1);
            double xg, yg;
            node_xy(dem, ip, &xg, &yg);
            double dx2 = (xx - xg), dy2 = (yy - yg);
            dx2 = dx2 * dx2;
            dy2 = dy2 * dy2;
            double dist = fmax(1.e-6, sqrt(dx2 + dy2));
            auxv[2 * ip] = auxv[2 * ip] + eb / dist;
            auxv[2 * ip + 1] = auxv[2 * ip + 1] + 1.0 / dist;
          }
        mass[i] = mass[i] + dt * o.er * mass[i] / rho[i];
        if (n == 1) rho[i] = rho[i] + dt * o.er; /* npoin_s = 1 (6:8756-8758) */
      }
      for (int64_t ip = 0; ip < npg; ip++) {
        if (auxv[2 * ip + 1] > 1.e-6) {
          double xnum = auxv[2 * ip] / auxv[2 * ip + 1];
          double *t = eros + 3 * ip; /* topol(1), topol(14), topol(15) */
          if (xnum > t[2]) xnum = t[2];
          if (t[2] > 0.) {
            t[1] = t[1] + xnum;
            t[0] = t[0] - xnum;
            t[2] = t[2] - xnum;
          }
        }
      }
    }
    /* 8. pore pressure on the column, npwp > 1 (Get_Pwp_SW 6:10966-11174) */
    if (ncol > 1) {
      pwp_column(p, &D, dem, x, n, ncol, dt, rho, Drho, flag, uaux);
      for (int i = 0; i < n; i++) u[i] = uaux[(int64_t)i * ncol];
    } else if (nuaux >= 1) { /* npwp = 1 (Get_Pwp_SW 6:11176-11310) */
      for (int i = 0; i < n; i++) {
        if (flag[i]) {
          uaux[i] = 0.0;
          continue;
        }
        double hh = rho[i], dhh = Drho[i], hinv;
        if (hh < D.comp) {
          hh = D.comp;
          hinv = 1.0 / D.comp;
        } else {
          hinv = 1 / hh;
        }
        double betaw = D.Bfact * hinv * hinv / D.alphaS2;
        uaux[i] = uaux[i] * exp(-betaw * dt) + (dhh / dt) * (1. - exp(-betaw * dt)) / betaw;
        if (uaux[i] > hh) uaux[i] = hh;
        if (uaux[i] < hh * D.alphaS * D.fpwp_min) uaux[i] = hh * D.alphaS * D.fpwp_min;
      }
      for (int i = 0; i < n; i++) u[i] = uaux[i];
    } else {
      for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        if (u[i] > rho[i]) u[i] = rho[i];
        if (u[i] < 0.0) u[i] = 0.0;
      }
    }
    /* 10. h_upgrade (7:2866-2961) then XSPH (7:3042-3099, 7:987-993) */
    if (p->sle == 1) { /* 7:2925-2931 */
      for (int i = 0; i < n; i++)
        if (!flag[i]) hsml[i] = 2.0 * pow(mass[i] / rho[i], 1. / 2);
    } else if (p->sle == 2) { /* dh/dt = -(h/dim) (1/rho) drho/dt, 7:2893-2919 */
      double *vcc = (double *)calloc(n, sizeof(double));
      for (int64_t kk = 0; kk < pv.n; kk++) {
        int64_t k = p->rev_order ? pv.n - 1 - kk : kk;
        int i = pv.i[k], j = pv.j[k];
        double dvx0 = v[2 * j] - v[2 * i], dvx1 = v[2 * j + 1] - v[2 * i + 1];
        double hvcc = dvx0 * pdw[2 * k];
        hvcc = hvcc + dvx1 * pdw[2 * k + 1];
        vcc[i] = vcc[i] + mass[j] * hvcc / rho[j];
        vcc[j] = vcc[j] + mass[i] * hvcc / rho[i];
      }
      for (int i = 0; i < n; i++) {
        if (flag[i]) continue;
        double dh = (hsml[i] / 2) * vcc[i];
        hsml[i] = hsml[i] + dt * dh;
        if (hsml[i] <= 0) hsml[i] = hsml[i] - dt * dh;
      }
      free(vcc);
    } else if (p->sle == 3) { /* McDougall & Hungr: hsml = B / sqrt(mean h/m), 7:2933-2959 */
      int *nnb = (int *)calloc(n, sizeof(int));
      double *hbm = (double *)calloc(n, sizeof(double));
      for (int64_t kk = 0; kk < pv.n; kk++) {
        int64_t k = p->rev_order ? pv.n - 1 - kk : kk;
        int i = pv.i[k], j = pv.j[k];
        nnb[i]++;
        nnb[j]++;
        hbm[i] = hbm[i] + rho[i] / mass[i];
        hbm[j] = hbm[j] + rho[j] / mass[j];
      }
      for (int i = 0; i < n; i++) {
        /* no neighbours: 0/0 in the reference (a NaN smoothing length);
           kept unchanged here */
        if (flag[i] || nnb[i] == 0) continue;
        double Ls = sqrt(hbm[i] / nnb[i]);
        hsml[i] = 4 / Ls;
      }
      free(nnb);
      free(hbm);
    }
    if (p->average_velocity) {
      memset(av, 0, sizeof(double) * 2 * n);
      for (int64_t kk = 0; kk < pv.n; kk++) {
        int64_t k = p->rev_order ? pv.n - 1 - kk : kk;
        int i = pv.i[k], j = pv.j[k];
        if (!(i >= p->npoin0 && j >= p->npoin0)) continue;
        for (int d = 0; d < 2; d++) {
          double dvx = v[2 * i + d] - v[2 * j + d];
          av[2 * i + d] = av[2 * i + d] - 2 * mass[j] * dvx / (rho[i] + rho[j]) * pw[k];
          av[2 * j + d] = av[2 * j + d] + 2 * mass[i] * dvx / (rho[i] + rho[j]) * pw[k];
        }
      }
      for (int i = p->npoin0; i < n; i++) {
        if (flag[i]) continue;
        v[2 * i] = v[2 * i] + 0.3 * av[2 * i];
        v[2 * i + 1] = v[2 * i + 1] + 0.3 * av[2 * i + 1];
      }
    }
    free(rhoF);
    /* time_integration_sph: Check_Out_Domain after RK45 (7:268) */
    for (int i = 0; i < n; i++)
      if (out_of_domain(dem, x[2 * i], x[2 * i + 1])) flag[i] = 1;
    time_sph = time_sph + dt; /* 8:1617 */
  }
  free(x0);
  free(v0);
  free(u0);
  free(rho0);
  free(S);
  free(inf);
  free(ex);
  free(du);
  free(Dx);
  free(Dv);
  free(Du);
  free(p_);
  free(av);
  free(divv);
  free(pw);
  free(pdw);
  free(pv.i);
  free(pv.j);
  free(auxv);
  if (tau0_io) *tau0_io = tau0_carry;
  if (dt_time_io) {
    dt_time_io[0] = dt;
    dt_time_io[1] = time_sph;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* a14 (i). OutputRes_WIR IDW scatter (6:14312-14466), soil only.       */
/* out[npoig][4] = h, vx, vy, u/h  (normalised where sum 1/d >= 1e-6)   */
/* wsum[npoig] = sum 1/d                                                */
/* ------------------------------------------------------------------ */
void gfo_raster_idw(const gfo_dem *d, int n, const double *x, const double *v, const double *rho,
                    const double *u, const int32_t *flag, double *out, double *wsum) {
  int64_t np = (int64_t)d->nx * d->ny;
  memset(out, 0, sizeof(double) * 4 * np);
  memset(wsum, 0, sizeof(double) * np);
  double distg = sqrt(d->dx * d->dx + d->dy * d->dy);
  for (int i = 0; i < n; i++) {
    if (flag[i]) continue;
    double xx = x[2 * i], yy = x[2 * i + 1];
    int ipx = (int)((xx - d->xmin) / d->dx + 1), ipy = (int)((yy - d->ymin) / d->dy + 1);
    int inx0 = ipx - 1 > 1 ? ipx - 1 : 1, iny0 = ipy - 1 > 1 ? ipy - 1 : 1;
    int inx1 = ipx + 2 < d->nx ? ipx + 2 : d->nx, iny1 = ipy + 2 < d->ny ? ipy + 2 : d->ny;
    for (int inx = inx0; inx <= inx1; inx++)
      for (int iny = iny0; iny <= iny1; iny++) {
        int64_t g = (int64_t)(iny - 1) * d->nx + inx - 1;
        double gx = d->xmin + (inx - 1) * d->dx, gy = d->ymin + (iny - 1) * d->dy;
        if (d->coorg) {
          gx = d->coorg[2 * g];
          gy = d->coorg[2 * g + 1];
        }
        double a = xx - gx, b = yy - gy;
        a = a * a;
        b = b * b;
        double dist = fmax(1.e-6, sqrt(a + b));
        dist = dist / distg;
        out[4 * g + 0] += rho[i] / dist;
        out[4 * g + 1] += v[2 * i] / dist;
        out[4 * g + 2] += v[2 * i + 1] / dist;
        out[4 * g + 3] += (u[i] / rho[i]) / dist;
        wsum[g] += 1.00 / dist;
      }
  }
  for (int64_t g = 0; g < np; g++)
    if (wsum[g] >= 1.e-6)
      for (int k = 0; k < 4; k++) out[4 * g + k] /= wsum[g];
}

/* ------------------------------------------------------------------ */
/* a14 (ii). Get_MC_pixel_Vars_SW (6:20717-20833): per plot time each   */
/* active particle adds to the 4 corners of its DEM cell.               */
/* acc[npoig][4] = sum h, max h, sum |v|, max |v| ; cnt[npoig] visits   */
/* ------------------------------------------------------------------ */
void gfo_mc_pixel_accum(const gfo_dem *d, int n, const double *x, const double *v,
                        const double *rho, const int32_t *flag, double *acc, int32_t *cnt) {
  double xmax = d->xmax_set ? d->xmax : d->xmin + (d->nx - 1) * d->dx;
  double ymax = d->xmax_set ? d->ymax : d->ymin + (d->ny - 1) * d->dy;
  for (int i = 0; i < n; i++) {
    if (flag[i]) continue;
    double xx = x[2 * i], yy = x[2 * i + 1];
    int ipx = (int)((xx - d->xmin) / d->dx + 1), ipy = (int)((yy - d->ymin) / d->dy + 1);
    if (ipx == d->nx && fabs(xmax - xx) <= 1.e-3) ipx--;
    if (ipy == d->ny && fabs(ymax - yy) <= 1.e-3) ipy--;
    if (ipx < 1 || ipy < 1 || ipx > d->nx - 1 || ipy > d->ny - 1) continue;
    int64_t g1 = (int64_t)(ipy - 1) * d->nx + ipx - 1;
    int64_t nodes[4] = {g1, g1 + 1, g1 + 1 + d->nx, g1 + d->nx};
    double hh = rho[i];
    double vv = sqrt(v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1]);
    for (int k = 0; k < 4; k++) {
      int64_t g = nodes[k];
      cnt[g] += 1;
      acc[4 * g + 0] += hh;
      acc[4 * g + 2] += vv;
      if (hh > acc[4 * g + 1]) acc[4 * g + 1] = hh;
      if (vv > acc[4 * g + 3]) acc[4 * g + 3] = vv;
    }
  }
}

/* ------------------------------------------------------------------ */
/* a14 (iii). Norm_MC_pixel_Vars_SW (6:20838-20908), called with an      */
/* offset after every problem (7:3521, 8:1539-1541): per node reached    */
/* by a problem (cnt > 0), the time averages OCMC(1), OCMC(3) = sums /   */
/* visits (6:20857-20860) and the two maxima enter MC_Pr(1,:) += OCMC,   */
/* MC_Pr(2,:) += OCMC**2 (6:20864-20869), nr_problem_pix += 1; after the */
/* last problem MC_Pr(3,:) = mean = MC_Pr(1,:)/nr, MC_Pr(4,:) =          */
/* sqrt(MC_Pr(2,:)/nr - mean**2), 0 below 1e-6 (6:20885-20898).         */
/* acc[R][npoig][4], cnt[R][npoig] as gfo_mc_pixel_accum leaves them.   */
/* Outputs: pr[npoig][4][4] = MC_Pr(row, var) for rows 1..4 as          */
/* pr[g][row][var]; nprob[npoig]; maxs[npoig][2] = max over problems of */
/* max h, max |v|; pexc[npoig] = fraction of problems with max h > hthr */
/* (the exceedance raster, no reference counterpart).                   */
/* ------------------------------------------------------------------ */
void gfo_mc_hazard(int64_t npoig, int R, const double *acc, const int32_t *cnt, double hthr, double *pr,
                   int32_t *nprob, double *maxs, double *pexc) {
  for (int64_t g = 0; g < npoig; g++) {
    double *P = pr + 16 * g;
    for (int k = 0; k < 16; k++) P[k] = 0.0;
    int nr = 0, nexc = 0;
    double mh = 0.0, mv = 0.0;
    for (int r = 0; r < R; r++) {
      const int64_t k = (int64_t)r * npoig + g;
      const int c = cnt[k];
      if (c == 0) continue;
      const double *a = acc + 4 * k;
      double o[4];
      o[0] = a[0] / c;
      o[1] = a[1];
      o[2] = a[2] / c;
      o[3] = a[3];
      for (int q = 0; q < 4; q++) {
        P[q] = P[q] + o[q];
        P[4 + q] = P[4 + q] + o[q] * o[q];
      }
      nr++;
      if (a[1] > mh) mh = a[1];
      if (a[3] > mv) mv = a[3];
      if (a[1] > hthr) nexc++;
    }
    const int nd = nr > 0 ? nr : 1;
    for (int q = 0; q < 4; q++) {
      const double mean = P[q] / nd, sig2 = P[4 + q] / nd;
      double t = sig2 - mean * mean;
      if (t < 1.e-6) t = 0.0;
      P[8 + q] = mean;
      P[12 + q] = sqrt(t);
    }
    nprob[g] = nr;
    maxs[2 * g] = mh;
    maxs[2 * g + 1] = mv;
    pexc[g] = R > 0 ? (double)nexc / R : 0.0;
  }
}
