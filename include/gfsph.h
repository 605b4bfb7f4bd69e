/*
 * gfsph.h -- C ABI of the MI355X-native Geoflow-SPH propagation step.
 *
 * One FP64 depth-integrated SPH time step on one GPU, as the reference's
 * time_integration_sph / RK45 perform it ("7 SPH_MAIN_2019.f90":220-283,
 * 746-1005).  Plain pointers and sizes only: a Fortran driver binds it with
 * ISO_C_BINDING, Python with ctypes.
 *
 * Array conventions.  Particle arrays are in the caller's particle order.
 * Two-component arrays are [n][2] (x then y), which is the memory layout of
 * the reference's Fortran x(ndimn, npoin) / vx(ndimn, npoin) arrays
 * ("5 SPH_Main_Vars_2019.f90":29-35), so they pass unchanged.  Pair indices
 * are 0-based and i < j.
 *
 * Every entry point returns GFSPH_OK (0) or a negative status; the message of
 * the last failure on a context is returned by gfsph_last_error().  A context
 * is used by one host thread at a time.
 */
#ifndef GFSPH_H
#define GFSPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFSPH_OK 0
#define GFSPH_ERR_ARG (-1)         /* invalid argument or inconsistent sizes      */
#define GFSPH_ERR_UNSUPPORTED (-2) /* option outside the implemented path         */
#define GFSPH_ERR_HIP (-3)         /* HIP runtime failure                         */
#define GFSPH_ERR_NONFINITE (-4)   /* NaN/Inf detected in the particle state      */
#define GFSPH_ERR_CAPACITY (-5)    /* output buffer too small (size reported)     */

typedef struct gfsph_ctx gfsph_ctx;

/* Solver switches and material constants.  Replaces the module state that
 * the reference reads from SPH_Main_Vars_2019 / SPH_SW_Vars_2019 / the time
 * variables ("5 SPH_Main_Vars_2019.f90":85-122, "4 SPH_SW_Vars_2019.f90":27,
 * "3 SPH_FEM_Driver_time_vars_2019.f90":26-37).  cst[k] is const(k),
 * k = 1..110, holding the REAL(4) value the reference reads (widened). */
typedef struct gfsph_params {
  int32_t integrator;        /* 4: RK4 (SPH_t_Integ_Alg=4); 2: RK2 (Heun)     */
  int32_t pa_sph;            /* 1, 2 or 4                                      */
  int32_t nnps;              /* accepted: 1..4 (all give the same pair set)    */
  int32_t sle;               /* 0 or 1                                         */
  int32_t skf;               /* 1 cubic spline, 2 Gauss, 3 quintic, 4 Wendland */
  int32_t summation_density; /* must be 1                                      */
  int32_t average_velocity;  /* XSPH on/off                                    */
  int32_t nor_density;       /* must be 0                                      */
  int32_t icpwp;             /* 0, 1 or 3                                      */
  int32_t npoin0;            /* particles [0, npoin0) are skipped by XSPH      */
  int32_t init_input;        /* 1: apply Input_SW initialisation at create     */
  int32_t check_finite;      /* 1: gfsph_step fails with GFSPH_ERR_NONFINITE   */
  double dt;                 /* dt_sph                                         */
  double h_inf;              /* h_inf_SW                                       */
  double pwp0rel;            /* initial pore-pressure ratio                    */
  double cst[111];           /* const(1..110), index 0 unused                  */
  /* icpwp = 2: pore pressure on a column of nuaux >= 3 nodes per particle,
   * bed to surface (Uaux(1:nUaux,:), Initialize_Pwp "6 SPH_SW_2019.f90":
   * 1743-1800, Get_Pwp_SW 6:10966-11174).  Ignored for other icpwp. */
  int32_t nuaux;
  int32_t pwp_bc_type[2];    /* BC_pwp_type bed, surface: 1 value, 2 gradient  */
  int32_t pwp0_type;         /* initial profile 0..4 (init_input = 1)          */
  double pwp_bc_val[2];      /* BC_pwp_val (relative to the depth)             */
  double xirel;              /* pwp0_type = 4: relative height of the profile  */
  /* ic_K0 = 1 (read by the reference for frictional laws only,
   * 6:1373-1390): the pressure times k0_active where the velocity divergence
   * is >= 0, else k0_pasive (6:6736-6760, 6:6857-6863).  With it, or with
   * ic_cosTheta = const(16) = 1, the pressure force is re-evaluated in every
   * RK stage, as IntForces_SW is. */
  int32_t ic_k0;
  int32_t pad_k0;
  double k0_active, k0_pasive;
} gfsph_params;

/* One terrain zone of the terrain file's ic_topo_props = 10 block
 * ("2 SPH_FEM_topo_2019.f90":516-576: basal_type, basal_law, basal_props).
 * type 1: basal friction -- a particle of a frictional law inside the zone
 * takes tan(phi) = props[6] when props[6] >= 0 (Get_Topo_Basal_Friction_SW,
 * "6 SPH_SW_2019.f90":10515-10545, applied at 6:9697-9702).  type 2: erosion
 * -- inside the zone the zone's own law replaces the erosion rate
 * (Get_Erosion_Rate_SW 6:10407-10505): law 0 none, 1 props[6] h |v|, 2 the
 * slope law with K = props[6], C* = props[7], 3 the basal-stress law with
 * K, c, tan(phi_b), beta = props[6..9].  type 3: pore pressure -- with a
 * column (icpwp = 2) the bed condition of a particle inside the zone at the
 * step's end is type int(props[6]), value props[7] (Get_Pwp_SW
 * 6:10987-11001).  props[0] is the geometry
 * (Which_Topo_Zone 2:2635-2808): 1, 2, 3 a range props[1]..props[2] of the
 * terrain Z, of x, of y; 4 the rectangle x props[1]..props[2], y
 * props[3]..props[4]; 5 the rectangle of corner (props[1], props[2]), sides
 * props[3], props[4] turned props[5] degrees; 101 the cells touching a node
 * of the zone's node list (gfsph_dem.zone_nodes).  The first zone of the
 * type that holds the point is taken. */
typedef struct gfsph_zone {
  int32_t type;      /* basal_type: 1 friction, 2 erosion, 3 pore pressure   */
  int32_t law;       /* basal_law: the erosion law of a type-2 zone (0..3)    */
  double props[16];  /* basal_props(1:16)                                     */
} gfsph_zone;

/* Regular DEM grid ("2 SPH_FEM_topo_2019.f90":94-110, 223-252).  Z is
 * [ny][nx] (node ipoig = iy*nx+ix).  If topol is NULL the nodal slopes and
 * curvature terms are derived on the device by central differences
 * (Fdiff_derivs, 2:2812-2877); otherwise topol is [ny*nx][9] holding
 * Z, Zx, Zy, 1+Zx^2, ZxZy, 1+Zy^2, Zxx/rt, Zxy/rt, Zyy/rt.
 * Terrain zones (all zero: none): nzones <= 32 zones of types 1 and 2
 * (gfsph_zone); zone_nodes[ipoig] has bit z set when the node is in the node
 * list of zone z (geometry 101: the nodes Get_Eros_Pts, 2:909-1110, gave an
 * erodible depth); eros_avail is topol(15,:) as the terrain reading leaves
 * it (those depths, 0 elsewhere), the available depth of zone erosion when
 * const(21) = 0 (Input_SW 6:1288-1308; with const(21) > 0 the reference
 * sets every node to const(48) or 999). */
typedef struct gfsph_dem {
  int32_t nx, ny;
  double xmin, ymin, dx, dy; /* Xming, Yming, deltxg, deltyg                   */
  double xmax, ymax;         /* Xmaxg, Ymaxg as read                           */
  const double *Z;           /* [ny*nx]                                        */
  const int32_t *in_model;   /* [ny*nx] 0/1 or NULL (all in model)             */
  const double *topol;       /* [ny*nx][9] or NULL                             */
  int32_t nzones;            /* ntopo_zones (ic_topo_props = 10), 0: none      */
  int32_t pad_z;
  const gfsph_zone *zones;   /* [nzones]                                       */
  const int32_t *zone_nodes; /* [ny*nx] node-list bits, or NULL                */
  const double *eros_avail;  /* [ny*nx] topol(15,:) of the terrain, or NULL    */
} gfsph_dem;

/* Uploads the DEM and the particle set and, if p->init_input, performs the
 * Input_SW initialisation: summed depth from the initial positions and pore
 * pressure u = pwp0rel*h*const(14) ("6 SPH_SW_2019.f90":1428-1452,
 * 1705-1852).  Replaces Init_sph ("7 SPH_MAIN_2019.f90":74) and
 * Setup_Global_Arrays_SW ("6 SPH_SW_2019.f90":6223).  itype must be all 1
 * (single-phase soil). */
int gfsph_create(const gfsph_params *p, const gfsph_dem *dem, int32_t n, const double *x,
                 const double *v, const double *h, const double *mass, const double *hsml,
                 const double *u, const int32_t *itype, gfsph_ctx **out);

/* nsteps calls of time_integration_sph ("7 SPH_MAIN_2019.f90":220-283),
 * i.e. the Driver_Flow step loop body ("8 SPH_FEM_Driver_2019.f90":1614-1620)
 * with fixed dt.  Asynchronous unless check_finite is set. */
int gfsph_step(gfsph_ctx *ctx, int32_t nsteps);

/* Waits for all work queued on the context's stream. */
int gfsph_sync(gfsph_ctx *ctx);

/* Copies the state back in the caller's particle order (any pointer may be
 * NULL).  h is the depth ("rho"), uaux the bottom pore pressure Uaux(1,:). */
int gfsph_get_state(gfsph_ctx *ctx, double *x, double *v, double *h, double *u, double *uaux,
                    double *hsml, int32_t *out_flag);

/* icpwp = 2: the pore-pressure columns [n][nuaux] (Uaux(:, ipoin), node 1
 * at the bed) in the caller's order; set replaces them, e.g. with the
 * columns the reference's Initialize_Pwp or a restart produced (u and the
 * bed value follow node 1). */
int gfsph_get_pwp_column(gfsph_ctx *ctx, double *col);
int gfsph_set_pwp_column(gfsph_ctx *ctx, const double *col);

/* Particle masses in the caller's order.  Erosion (below) changes them:
 * mass(:) after ExtForces_SW(1) ("6 SPH_SW_2019.f90":8755). */
int gfsph_get_mass(gfsph_ctx *ctx, double *mass);

/* mass0(:) ("5 SPH_Main_Vars_2019.f90":41), the masses before any erosion,
 * in the caller's order: the Bingham-erosion law nfric = 3 reads (mass -
 * mass0)/mass ("6 SPH_SW_2019.f90":9451).  gfsph_create takes the masses it
 * is given, as Setup_Global_Arrays_SW does (6:6398); a caller that creates
 * the context from a state that has already eroded hands the reference's
 * mass0 over here.  GFSPH_ERR_ARG unless the context has nfric = 3. */
int gfsph_set_mass0(gfsph_ctx *ctx, const double *mass0);

/* Erosion.  With const(21) > 0 (law 1 Hungr, 2 slope angle, 3 basal
 * stress; or the older form const(4) > 1e-4) every RK stage adds the
 * erosion momentum term and every step ends with ExtForces_SW(1): the
 * particles gain mass at the erosion rate and the terrain loses the eroded
 * depth ("6 SPH_SW_2019.f90":8664-8812, Get_Erosion_Rate_SW 10289-10404).
 * The terrain's erosion state of realisation r (0 for a single run) is three
 * [ny*nx] arrays: the current Z (topol(1,:)), the eroded depth
 * (topol(14,:)) and the depth still available (topol(15,:)).  At create Z is
 * dem->Z, the eroded depth 0 and the available depth const(48) if given,
 * else 999 (6:1288-1308).  gfsph_erosion_set replaces any of them (NULL:
 * keep), as Get_Topo_restart (2:2880-2921) reads them from restart.top.
 * Both fail with GFSPH_ERR_ARG when erosion is off. */
int gfsph_erosion_read(gfsph_ctx *ctx, int32_t r, double *z, double *eroded, double *avail);
int gfsph_erosion_set(gfsph_ctx *ctx, int32_t r, const double *z, const double *eroded, const double *avail);

/* Neighbour set of grid_find_NEW_tools3 ("5 a My_SPH_Tools_2019.f90":802-1104)
 * for the given positions: pairs (i<j) with r < scale_k*(hsml_i+hsml_j)/2.
 * Stateless; used by the parity tests and the Pint_Update export.  On
 * GFSPH_ERR_CAPACITY *npairs holds the required size. */
int gfsph_neighbour_pairs(int32_t skf, int32_t n, const double *x, const double *hsml,
                          const int32_t *flag, int64_t cap, int64_t *npairs, int32_t *ij);

/* The neighbour set of gfsph_neighbour_pairs without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  x [n][2], hsml [n], flag [n] (NULL: all active)
 * and the outputs are caller-owned DEVICE memory; the work is enqueued on
 * stream (a hipStream_t; NULL: the null stream) and nothing is waited for.
 * *npairs (device) receives the number of pairs (i < j, caller indices);
 * the first min(*npairs, cap) are written to ij [cap][2], in no particular
 * order.  Every intermediate buffer lives in one static device array of
 * about 13 MB per GPU, held until the process ends: n is limited to
 * 262,144 (GFSPH_ERR_CAPACITY beyond, with a message in
 * gfsph_last_error(NULL)), and calls that overlap on one GPU race. */
int gfsph_neighbour_pairs_noalloc(int32_t skf, int32_t n, const double *x, const double *hsml,
                                  const int32_t *flag, int64_t cap, uint64_t *npairs, int32_t *ij, void *stream);

/* Particle-to-DEM-node inverse-distance raster of OutputRes_WIR
 * ("6 SPH_SW_2019.f90":14312-14466): out[ny*nx][4] = h, vx, vy, u/h,
 * wsum[ny*nx] = sum of 1/d. */
int gfsph_raster_idw(gfsph_ctx *ctx, double *out, double *wsum);

/* Monte Carlo pixel accumulation of Get_MC_pixel_Vars_SW
 * ("6 SPH_SW_2019.f90":20717-20833) into the context's device rasters:
 * sum h, max h, sum |v|, max |v| and the visit count of the 4 nodes of each
 * particle's DEM cell.  gfsph_mc_pixel_read copies them out and, if reset,
 * clears them. */
int gfsph_mc_pixel_accumulate(gfsph_ctx *ctx);
int gfsph_mc_pixel_read(gfsph_ctx *ctx, double *acc /*[npoig][4]*/, int32_t *cnt, int32_t reset);

/* Batched Monte Carlo: nreal independent realisations (the ic_MCarlo loop of
 * Driver_Input / Driver_Flow, "8 SPH_FEM_Driver_2019.f90":1282-1439 and
 * 1554-1710, run together instead of one after another) on one DEM, stepped
 * by gfsph_step in one launch sequence.  Realisation r has counts[r]
 * particles, stored one realisation after another in x, v, h, mass, hsml, u
 * (n = sum of counts), and its own const(:) in cst[r*111 .. r*111+110]
 * (cst == NULL: p->cst for all; Update_Const_SW, 6:20035).  The switches of
 * p are shared, and so must be const(1) and const(14).  Realisations never
 * interact; each advances exactly as gfsph_create + gfsph_step would advance
 * it alone.  Bingham (nfric = 4): either every realisation or none; each
 * carries its own Tau0 from particle to particle and call to call, as a run
 * alone does (a batch mixing Bingham with other laws: GFSPH_ERR_UNSUPPORTED).
 * gfsph_get_state returns all particles in the same stacked order. */
int gfsph_batch_create(const gfsph_params *p, const gfsph_dem *dem, int32_t nreal, const int32_t *counts,
                       const double *cst, const double *x, const double *v, const double *h,
                       const double *mass, const double *hsml, const double *u, gfsph_ctx **out);
/* Get_MC_pixel_Vars_SW ("6 SPH_SW_2019.f90":20717-20833) for every
 * realisation at the current (plot) time, into per-realisation node
 * accumulators [nreal][npoig] (sum h, max h, sum |v|, max |v|, visits),
 * allocated on first use.  gfsph_batch_pixel_read copies realisation r's
 * accumulators out. */
int gfsph_batch_pixel_accumulate(gfsph_ctx *ctx);
int gfsph_batch_pixel_read(gfsph_ctx *ctx, int32_t r, double *acc /*[npoig][4]*/, int32_t *cnt /*[npoig]*/);
/* Per-node statistics over the realisations (Norm_MC_pixel_Vars_SW,
 * "6 SPH_SW_2019.f90":20838-20908): stats[npoig][4][2] = mean and standard
 * deviation, over the nprob[node] realisations that reached the node, of
 * (average h, max h, average |v|, max |v|); maxs[npoig][2] = max h and max |v|
 * over all realisations; pexc[npoig] = fraction of the realisations whose
 * max h exceeds hthr (the exceedance-probability raster, a build extension
 * with no reference counterpart). */
int gfsph_batch_hazard(gfsph_ctx *ctx, double hthr, double *stats, int32_t *nprob, double *maxs, double *pexc);
/* Picks the work-group size (64..512) of each neighbour-gather kernel
 * (density, pressure force, XSPH) by timing reps launches of each on the
 * context's current state; no reference counterpart (the reference is
 * serial).  Results do not depend on the choice, and the state is not
 * advanced.  Needs one completed gfsph_step.  chosen[3] (may be NULL)
 * receives the sizes. */
int gfsph_autotune(gfsph_ctx *ctx, int32_t reps, int32_t *chosen);
/* Reverse mode (adjoint) of the step, for back-analysis of rheology.  No
 * reference counterpart: the reference has no derivatives.  Covered, for
 * one run or a batch whose realisations share the law: nfric = 5
 * (frictional with Voellmy turbulence, xi = const(3) >= 0) and nfric = 4
 * (Bingham, its sticky Tau0 included: the cotangent of a carried Tau0 goes
 * to the particle, stage or step that set it), with pa_sph = 2, skf = 1,
 * icpwp = 0, sle = 0 (else GFSPH_ERR_UNSUPPORTED).
 *
 * gfsph_adjoint_forward advances nsteps steps from the current state like
 * gfsph_step, keeping a copy of the device state every ckpt_every steps.
 * gfsph_adjoint_backward then returns the vector-Jacobian product of those
 * nsteps steps: given the cotangents of the final state lam_x [n][2], lam_v
 * [n][2], lam_h [n] (caller order; NULL = 0), it recomputes each segment
 * from its checkpoint with each step's neighbour set fixed as in the forward
 * pass and sweeps backwards, giving, per realisation ([R]; R = 1 for a
 * single run), the cotangents of the law's two parameters -- g_mu, g_xi =
 * dL/dmu (mu = const(9)), dL/dxi (const(3)) for nfric 5; dL/dtau_y
 * (const(6)), dL/dmu_B (const(7)) for Bingham -- and dL/dmass [n], dL/dx0
 * [n][2], dL/dv0 [n][2] at the state the forward pass started from (any
 * output may be NULL).  The context is left in the state the forward pass
 * ended in.  Particles that leave the domain inside a step contribute no
 * derivative for that step. */
int gfsph_adjoint_forward(gfsph_ctx *ctx, int32_t nsteps, int32_t ckpt_every);
/* Records at most tape_steps steps at once in gfsph_adjoint_backward (0: a
 * whole checkpoint segment).  A longer segment is split from its checkpoint
 * into at most 8 pieces, each with a checkpoint of its own, recursively,
 * until the pieces fit: every further level costs one more forward pass and
 * 8 checkpoints, the records of tape_steps steps stay the largest buffer --
 * the form of recomputation for horizons whose records do not fit beside the
 * checkpoints.  The gradients do not depend on it (every recomputed step
 * equals the forward's).  No reference counterpart. */
int gfsph_adjoint_set_tape(gfsph_ctx *ctx, int32_t tape_steps);
int gfsph_adjoint_backward(gfsph_ctx *ctx, const double *lam_x, const double *lam_v, const double *lam_h,
                           double *g_mu, double *g_xi, double *g_mass, double *g_x0, double *g_v0);
/* Deposit misfit of the current state: L = sum over DEM nodes of (R - T)^2,
 * R the IDW depth raster of OutputRes_WIR ("6 SPH_SW_2019.f90":14312-14466)
 * and T = target [npoig]; in a batch each realisation's raster against the
 * same target, loss [R]; lam_x [n][2] and lam_h [n] (may be NULL) receive
 * dL/dx and dL/dh in caller order.  No reference counterpart. */
int gfsph_raster_misfit(gfsph_ctx *ctx, const double *target, double *loss, double *lam_x, double *lam_h);
/* Average device time per launch of each step kernel over nrep steps, in
 * milliseconds, measured with HIP events on the context's stream.
 * names/ms are arrays of *nkernels entries (at most cap). */
int gfsph_profile_step(gfsph_ctx *ctx, int32_t nrep, int32_t cap, int32_t *nkernels,
                       const char **names, double *ms);

/* Number of particles and of active (in-domain) particles. */
int gfsph_counts(gfsph_ctx *ctx, int32_t *n, int32_t *nactive);

/* Time-step control ic_adapt_dt_sph: replaces adaptive_dt_sph /
 * Adaptive_dt_sph_SW ("7 SPH_MAIN_2019.f90":289-306, "6 SPH_SW_2019.f90":
 * 6521-6664), which Driver_Flow calls before each step ("8 SPH_FEM_Driver_2019.f90":
 * 1590).  mode 0 keeps dt_sph fixed; 1: min hsml/(3c), 2: min hsml/(3(c+|v|)),
 * 4: min |dx_DEM|/|v|, 10: min of 1 and 4 (c = sqrt(const(1) h), h clamped
 * to h_inf in place), 7: min 0.5 |x_(k+1) - x_k| / (|vx_(k+1) + vx_k| / 2)
 * over consecutive particles of the caller's order (the one-dimensional free
 * surface), dt_sph growing by at most 10 % per step.  Every following step
 * first computes the new dt_sph of each realisation on the device.  Mode 3
 * gives GFSPH_ERR_UNSUPPORTED: it halves its running minimum at every pair
 * (6:6587-6590), so its dt_sph is 0 on any flow of more than ~1,100 pairs. */
int gfsph_set_adaptive_dt(gfsph_ctx *ctx, int32_t mode);

/* time_sph (the sum of the dt_sph of the steps taken, "8 SPH_FEM_Driver_2019.f90":
 * 1617) and the dt_sph of the last step, one value per realisation (either
 * pointer may be NULL).  Synchronises. */
int gfsph_time(gfsph_ctx *ctx, double *time, double *dt);

/* Case control: every realisation of the context runs the step loop the
 * reference runs for each case of a case table or Monte Carlo set
 * (Driver_Flow_23, "8 SPH_FEM_Driver_2019.f90":1444-1548), on the device and
 * independently: each step's dt_sph is cut so that the step ends on the next
 * multiple of time_plot (8:1481-1488) and restored after it; at each plot time
 * the realisation's pixels are accumulated (Get_MC_pixel_Vars_SW,
 * "6 SPH_SW_2019.f90":20717-20833, if pixels); after each step Stop_case_SW
 * (6:20114-20177: mean |vx| below vstop after 10 steps, more than a fifth of
 * the particles out of the domain, time_sph beyond timestop, dt below
 * dtimestop) or time >= time_end stops the realisation, which then no longer
 * changes while the others continue.  gfsph_time reports each realisation's
 * time_sph and dt_sph.  Resets the loop state (t_plot_reset = 0, itimestep =
 * 0, none stopped). */
typedef struct gfsph_case_ctl {
  double vstop, timestop, dtimestop; /* vstop_cases, timestop_cases, dtimestop_cases */
  double dt;                         /* the driver's dt (Driver_Input, 8:330)        */
  double time_end;                   /* time_end                                     */
  double time_plot;                  /* plot_step*dt; 0: no plot times               */
  int32_t pixels;                    /* 1: pixel statistics at plot times            */
  int32_t pad_;
} gfsph_case_ctl;
int gfsph_batch_cases(gfsph_ctx *ctx, const gfsph_case_ctl *ctl);
/* Per realisation: stopped (0/1) and loop iterations done (itimestep); the
 * number stopped.  Any pointer may be NULL.  Synchronises. */
int gfsph_batch_status(gfsph_ctx *ctx, int32_t *stopped, int32_t *nsteps, int32_t *nstopped);

/* Multiplies the initial pore pressure of realisation r (u and the whole
 * Uaux column of every particle) by factor[r]: Input_SW's case variable 401
 * ("6 SPH_SW_2019.f90":1453-1479), applied after the context is created
 * (gfsph_inputs_batch_create calls it for each 401 column).  GFSPH_ERR_ARG
 * with icpwp = 0, which the reference's branch does not reach. */
int gfsph_batch_scale_pwp(gfsph_ctx *ctx, const double *factor);

/* Sets time_sph of every realisation, e.g. to time_sph_restart
 * ("6 SPH_SW_2019.f90":3302) when a run continues from restart.dat. */
int gfsph_set_time(gfsph_ctx *ctx, double time);

/* ---------------------------------------------------------------------------
 * The reference's input files, replacing what Init_sph reads
 * ("7 SPH_MAIN_2019.f90":74).  In directory dir:
 *   <name>.MASTER.dat  run control (Driver_Input, "8 SPH_FEM_Driver_2019.f90":
 *                      229-340), optional: SPH only, SPH_t_Integ_Alg 4
 *   <name>.dat         switches, h_inf_SW, const(:), icpwp (Input_SW,
 *                      "6 SPH_SW_2019.f90":812-1570): a single-phase soil release
 *   <topo>.top         terrain (Topo_main, "2 SPH_FEM_topo_2019.f90":208-620):
 *                      ictop 0 (explicit grid) or 10, 11, 15, 16 (scattered DTM
 *                      gridded as Get_DTM does, 2:1155-1449); +100 selects
 *                      finite-difference slopes, otherwise Get_Topol's
 *                      finite-element projection (2:1723-1939) runs on the GPU
 *   <name>.pts         release points gridded as Read_Source_plus_New does
 *                      (icunk_s 7, 6:4999-5272), or restart.dat (icunk_s 200,
 *                      Get_Restart_Info_SW, 6:3268-3340)
 * topo = NULL reads <name>.top.  Other options of these files are rejected
 * with GFSPH_ERR_UNSUPPORTED.  On failure *out still holds the object for
 * gfsph_inputs_error unless it could not be allocated.
 * ------------------------------------------------------------------------- */
typedef struct gfsph_inputs gfsph_inputs;

typedef struct gfsph_run {       /* <name>.MASTER.dat (all zero if absent)      */
  int32_t present;
  int32_t t_integ_alg;           /* SPH_t_Integ_Alg                             */
  int32_t adapt_dt;              /* ic_adapt_dt_sph                             */
  int32_t maxtimestep, print_step, save_step, plot_step;
  int32_t pad_;
  double dt, time_end, dt_sph;
  double time_restart;           /* time_sph_restart of restart.dat (icunk 200) */
} gfsph_run;

typedef struct gfsph_grid {      /* the terrain grid (Topo_main's module state) */
  int32_t nx, ny;                /* npoigx, npoigy                              */
  double xmin, ymin, dx, dy;     /* Xming, Yming, deltxg, deltyg                */
  double xmax, ymax;             /* Xmaxg, Ymaxg                                */
  double xshift, yshift;         /* xshiftg, yshiftg (ictop 11, 15, 16)         */
  int32_t fdiff;                 /* 1: slopes by finite differences at create  */
  int32_t pad_;
} gfsph_grid;

int gfsph_inputs_read(const char *dir, const char *name, const char *topo, gfsph_inputs **out);
/* The host half of gfsph_inputs_read: the files, the grids' geometry and the
 * binning of the points, with no device work (gfsph_inputs_info then reports
 * the switches, run control, grid and, for a restart, the particles). */
int gfsph_inputs_parse(const char *dir, const char *name, const char *topo, gfsph_inputs **out);
/* The parsed switches (dt = dt_sph of MASTER.dat, else 0; init_input = 1),
 * run control, grid and number of release particles. */
int gfsph_inputs_info(gfsph_inputs *in, gfsph_params *p, gfsph_run *run, gfsph_grid *grid, int32_t *npart);
/* Z, poig_in_model ([ny*nx]) and topol(1:9,:) ([ny*nx][9]; not available with
 * finite-difference slopes). */
int gfsph_inputs_dem(gfsph_inputs *in, double *Z, int32_t *in_model, double *topol);
/* The release as read or gridded, before Input_SW's summed-depth and
 * pore-pressure initialisation (which gfsph_inputs_create applies). */
int gfsph_inputs_particles(gfsph_inputs *in, double *x, double *v, double *h, double *mass, double *hsml,
                           double *u, int32_t *itype);
/* gfsph_create on the inputs (init_input = 1), then ic_adapt_dt_sph of
 * MASTER.dat and time_sph_restart. */
int gfsph_inputs_create(gfsph_inputs *in, gfsph_ctx **out);
/* The case table of MASTER.dat (Driver_Input, "8 SPH_FEM_Driver_2019.f90":
 * 475-1236): ic_cases_mat 1 one variable at a time, 2 every combination, 3
 * n_MCarlo_problems rows drawn from the combinations, 4 / 5 / 31 / 32 values
 * drawn from each variable's law (constant, triangular, normal, log-normal,
 * histogram: GetMCarlo_Cont_values / GetMCarlo_Dscrt_values, 8:1282-1439),
 * 20 / 21 the 2 ncases + 1 FOSM sets (means, then every variable at mean +-
 * sigma/10 as the reference fills them, 8:974-1051).  table is OC_mat,
 * [nproblems][3 + ncases + 2] (case, value index, variable index, the value
 * of every case variable, the y shift of variable 521); stop3 =
 * vstop_cases, timestop_cases, dtimestop_cases.  Any pointer may be NULL
 * (query the sizes first).  The reference seeds its draws from the system
 * clock (8:1306-1312); here they come from a seed (1 after reading, see
 * gfsph_inputs_resample): variable k's uniforms are SplitMix64 from state
 * seed + 0x9E3779B97F4A7C15 (k + 1), output u giving ((u >> 11) + 0.5) 2^-53;
 * mode 3 draws its rows from stream k = 0.  Mode 5 is read as mode 4 with
 * one geological zone; mode 6 (the trigger module's FOSM) and n_GZones > 1
 * (draws per geological zone of the terrain) are rejected. */
int gfsph_inputs_cases(gfsph_inputs *in, int32_t *mode, int32_t *ncases, int32_t *nproblems, double *table,
                       double *stop3);
/* Draws a sampled set's problems again from another seed (ic_cases_mat 3, 4,
 * 31, 32; GFSPH_ERR_ARG for the deterministic tables). */
int gfsph_inputs_resample(gfsph_inputs *in, uint64_t seed);
/* Erodible points of a geometry-101 terrain zone: Get_Eros_Pts
 * ("2 SPH_FEM_topo_2019.f90":909-1110) on the grid (nx, ny, xmin, ymin,
 * dx, dy; node coordinates cx, cy [ny*nx], or NULL for xmin + ix dx,
 * ymin + iy dy).  pts [npts][3] are the file's x, y, erodible depth; every
 * node within max(2, dx_s/dx) nodes of a point (and inside the points'
 * bounding box) takes its inverse-distance weight.  Out: avail = topol(15,:)
 * (the weighted depth, 0 where no point reached), listed = the zone's node
 * list (nodes some point reached), box = basal_props(2:5) (the points'
 * x min, x max, y min, y max); Z is changed by type_eros -1 (Z + depth),
 * -2 (max(Z, depth), depth = the rise), -3 (Z - depth). */
/* The terrain zones read (ic_topo_props = 10): their number, the zones as
 * read (geometry-101 zones with basal_props(2:5) = their points' box), the
 * node-list bits and topol(15,:) [ny*nx] (zeros without such zones).  Any
 * pointer may be NULL. */
int gfsph_inputs_zones(gfsph_inputs *in, int32_t *nzones, gfsph_zone *zones, int32_t *zone_nodes,
                       double *eros_avail);
int gfsph_terrain_eros_pts(int32_t nx, int32_t ny, double xmin, double ymin, double dx, double dy, const double *cx,
                           const double *cy, int32_t npts, const double *pts, int32_t type_eros, double dx_s,
                           double *Z, double *avail, int32_t *listed, double *box);
/* All problems of the case table as one batch (replaces the problem loop of
 * the reference's main program, 8:60-130): realisation p is the release and
 * const(:) as Input_SW changes them for problem p (6:1040-1232: variables 501
 * depth factor, 511 expansion, 521 shift, const indices; 401 the initial
 * pore pressure's factor, 6:1453-1479), with ic_adapt_dt_sph and the case
 * control (gfsph_batch_cases) of the loop the reference runs for each
 * problem: Driver_Flow for mode 1, Driver_Flow_23 (steps cut at every
 * plot_step*dt; pixel statistics if pixels) for the others. */
int gfsph_inputs_batch_create(gfsph_inputs *in, int32_t pixels, gfsph_ctx **out);
const char *gfsph_inputs_error(gfsph_inputs *in);
void gfsph_inputs_destroy(gfsph_inputs *in);

/* Rasters on the terrain grid to a NetCDF classic file (64-bit offset
 * variant): dimensions x (nx), y (ny), coordinate variables x and y, and one
 * double variable (y, x) per raster with CF attributes (units, long_name,
 * _FillValue unless fill is NaN).  data[k] is [ny][nx]; names are NetCDF
 * names (a letter, then letters, digits, '_'); each raster below 4 GiB.  No
 * reference counterpart in this format: replaces the hazard-map text outputs
 * written after the Monte Carlo loop ("6 SPH_SW_2019.f90":21200-21560) for
 * the statistics of gfsph_batch_hazard. */
int gfsph_write_netcdf(const char *path, int32_t nx, int32_t ny, double xmin, double ymin, double dx, double dy,
                       int32_t nvar, const char *const *names, const char *const *units,
                       const char *const *long_names, const double *const *data, double fill, const char *title);

/* The message of the last failure on ctx; with ctx = NULL, of the calling
 * thread's last failed context-free call (gfsph_neighbour_pairs_noalloc). */
const char *gfsph_last_error(gfsph_ctx *ctx);
void gfsph_destroy(gfsph_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
