/* gfsph_oracle.h -- interface of the serial FP64 CPU restatement (oracle).
 * TEST INFRASTRUCTURE ONLY; see gfsph_oracle.c. */
#ifndef GFSPH_ORACLE_H
#define GFSPH_ORACLE_H
#include <stdint.h>

typedef struct {
  int32_t integrator;         /* 4 = RK4 (RK45, 7:746); 2 = RK2 Heun (extension) */
  int32_t pa_sph, skf, sle;   /* 5:85-102 */
  int32_t summation_density;  /* must be 1 */
  int32_t average_velocity;   /* XSPH, 7:987 */
  int32_t nor_density;        /* must be 0 */
  int32_t icpwp;              /* 0, 1 or 3 (6:1418) */
  int32_t npoin0;             /* particles [0,npoin0) skip XSPH */
  int32_t rev_order;          /* 1: every pair sum in reverse order (conditioning probe) */
  int32_t adapt_dt;           /* ic_adapt_dt_sph: 0 fixed dt, or 1, 2, 4, 7, 10 (6:6521-6664) */
  int32_t pad2_;
  double dt;                  /* dt_sph (initial value when adapt_dt != 0) */
  double h_inf;               /* h_inf_SW */
  double pwp0rel;             /* initial pwp ratio (6:1811-1852) */
  double cst[111];            /* const(1..110); index 0 unused; REAL(4) values */
  /* icpwp = 2: pore pressure on a column of nuaux nodes per particle
   * (Initialize_Pwp 6:1743-1800, Get_Pwp_SW 6:10966-11174); uaux is then
   * [n][nuaux] (Uaux(1:nuaux, ipoin)), node 1 at the bed */
  int32_t nuaux;
  int32_t pwp_bc_type[2];     /* BC_pwp_type: bottom, top (1 value, 2 gradient) */
  int32_t pwp0_type;          /* initial profile 0..4 */
  double pwp_bc_val[2];       /* BC_pwp_val */
  double xirel;               /* pwp0_type 4 */
  /* K0 (frictional laws, 6:1373-1390): pressure times K0_active where the
   * velocity divergence is >= 0, else K0_pasive (6:6736-6760, 6:6857-6863) */
  int32_t ic_k0, pad3_;
  double k0_active, k0_pasive;
  /* mass0 (5:41): the masses before any erosion, read by the Bingham-erosion
   * law nfric = 3 (6:9451); required for it, else may be NULL */
  const double *mass0;
} gfo_params;

/* one terrain zone: basal_type, basal_law, basal_props(1:16) (2:516-576) */
typedef struct {
  int32_t type, law;
  double props[16];
} gfo_zone;

typedef struct {
  int32_t nx, ny;             /* npoigx, npoigy */
  int32_t xmax_set, pad_;     /* 1: use xmax/ymax below (Xmaxg/Ymaxg as read) */
  double xmin, ymin, dx, dy;  /* Xming, Yming, deltxg, deltyg */
  double xmax, ymax;
  const double *topol;        /* [nx*ny][9]: Z Zx Zy top4..top9 */
  const int32_t *in_model;    /* [nx*ny] or NULL (all 1) */
  const double *coorg;        /* [nx*ny][2] or NULL (regular grid) */
  /* terrain zones (ic_topo_props = 10): friction (type 1) and erosion
   * (type 2) zones; zone_nodes [nx*ny] bit z: node in zone z's node list
   * (geometry 101), or NULL */
  int32_t nzones, pad2_;
  const gfo_zone *zones;
  const int32_t *zone_nodes;
} gfo_dem;

void gfo_kernel(int skf, double r, double dx, double dy, double hsml, double *w, double *dwx,
                double *dwy);
int64_t gfo_pairs(int n, int skf, const double *x, const double *hsml, const int32_t *flag,
                  int brute, int64_t cap, int32_t *ij, double *w, double *dwdx);
void gfo_topo_lookup(const gfo_dem *d, int n, const double *xy, double *out9);
/* Which_Topo_Zone (2:2635-2808) at n points for basal type btype: zone
 * index (0-based) or -1; zc: current Z [nx*ny] or NULL (topol's Z) */
void gfo_which_zone(const gfo_dem *d, const double *zc, int n, const double *xy, int btype, int32_t *out);
void gfo_fdiff_derivs(int nx, int ny, double dxg, double dyg, const int32_t *in_model,
                      double *topol);
int gfo_init(const gfo_params *p, const gfo_dem *dem, int n, const double *x, double *h,
             const double *mass, const double *hsml, double *u, double *uaux, int32_t *flag);
/* Adaptive_dt_sph_SW (6:6521-6664) alone: the next dt_sph from dt; clamps
 * rho to h_inf_SW in place where the mode reads it */
double gfo_adaptive_dt(const gfo_params *p, const gfo_dem *dem, int n, const double *x, const double *v,
                       double *rho, const double *hsml, const int32_t *flag, double dt);
/* mass is updated in place by erosion; eros is [nx*ny][3] = topol(1),
 * topol(14), topol(15) in/out, required when erosion is on (else -2) */
int gfo_run(const gfo_params *p, const gfo_dem *dem, int n, double *x, double *v, double *rho,
            double *mass, double *hsml, double *u, double *uaux, int32_t *flag,
            int nsteps, double *tau0_io /* Bingham Tau0 carry, in/out; NULL = 0 */,
            double *dt_time_io /* [dt_sph, time_sph] in/out; NULL: p->dt, 0 */,
            double *eros /* erosion terrain state or NULL */);
void gfo_raster_idw(const gfo_dem *d, int n, const double *x, const double *v, const double *rho,
                    const double *u, const int32_t *flag, double *out, double *wsum);
void gfo_mc_pixel_accum(const gfo_dem *d, int n, const double *x, const double *v,
                        const double *rho, const int32_t *flag, double *acc, int32_t *cnt);
/* Norm_MC_pixel_Vars_SW over R problems' accumulators: pr[npoig][4][4]
 * (rows: sum, sum of squares, mean, standard deviation of average h, max h,
 * average |v|, max |v|), nprob, maxs[npoig][2], pexc */
void gfo_mc_hazard(int64_t npoig, int R, const double *acc, const int32_t *cnt, double hthr, double *pr,
                   int32_t *nprob, double *maxs, double *pexc);
#endif
