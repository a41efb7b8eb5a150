/*
 * gazpart_gpu.h -- C-ABI of the MI355X (gfx950) Lagrangian ellipsoid step.
 *
 * The reference (GAZPART, Fortran/MPI) has no plugin API: its hot path is a
 * set of external subroutines with implicit interfaces plus module-global
 * state.  Each entry point below replaces one of them; a Fortran caller
 * binds them with an iso_c_binding interface block (see
 * gazpart_amd/fortran/gazpart_gpu_mod.f90).  Citations are to the
 * reference's source/ directory.
 *
 * Conventions
 *  - Every call returns an int status (GP_OK = 0); nothing calls exit/stop.
 *  - Host arrays are owned by the caller and never retained after return.
 *  - Field and particles stay resident in device memory between calls.
 *  - One context per GPU; a context is not re-entrant.
 *  - Particle indices at the boundary are the reference's: 1-based, in the
 *    order the particles were given to gp_set_particles.
 *  - state layout: double[13][np], components x y z u v w qa qb qc qd
 *    ox oy oz (omega in the body frame, as PART%OMEGAX..).
 *  - field layout: double[6][n+4][n+4][n] = UAIN(1:n,-1:n+2,-1:n+2) for
 *    u v w wx wy wz (allocate_arrays.f90:108-118).
 */
#ifndef GAZPART_GPU_H
#define GAZPART_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gp_ctx gp_ctx;

/* status codes; 1-6 mirror the reference's `stop` statements */
enum {
  GP_OK = 0,
  GP_ERR_NAN_DRAG = 1,        /* fluid_force_ellipsoid.f90:150-158 */
  GP_ERR_NAN_OMEGA = 2,       /* fluid_torque_ellipsoid.f90:178-188 */
  GP_ERR_KE_COLLISION = 3,    /* collision_response.f90:408-436 */
  GP_ERR_KE_WALL = 4,         /* wall_ellipsoid_rebound.f90:372-384 */
  GP_ERR_OUT_OF_GRID = 5,     /* collision_detection_connectivity.f90:81-89 */
  GP_ERR_NCLOSE_OVERFLOW = 6, /* NCLOSE > NCLOSEMAX (collision.f90:35-37) */
  GP_ERR_INTERP_RANGE = 7,    /* stencil outside UAIN (interp_lag3.f90:124) */
  GP_ERR_CELL_RULE = 8,       /* gp_params.cell_rule = 0 and a species' largest
                                 semi-axis exceeds half a collision cell: the
                                 reference rule DPMAX = 2 max(EMAJ)
                                 (collision_initiation.f90:50-75) would lose
                                 candidate pairs (oblate APR < 0.5) */
  GP_ERR_ARG = 10,
  GP_ERR_STATE = 11,
  GP_ERR_HIP = 12,
  GP_ERR_NOMEM = 13,
  GP_ERR_IO = 14              /* a file could not be opened, read or written,
                                 or its framing / header does not match */
};

enum { GP_PREC_FP64 = 0, GP_PREC_MIXED = 1 };

/* species (particle class x size) row layout, GP_SP_STRIDE doubles */
enum {
  GP_SP_A = 0, GP_SP_B, GP_SP_C,   /* semi-axes (body x, y, z) */
  GP_SP_PMASS, GP_SP_DEQ,          /* PMASS, dequiv (fluid_force_ellipsoid.f90:59,70) */
  GP_SP_IXX, GP_SP_IYY, GP_SP_IZZ, /* inertia / mass */
  GP_SP_GRAV, GP_SP_ECP, GP_SP_MUP,
  GP_SP_RHOP, GP_SP_EMAJ, GP_SP_APR,
  GP_SP_STRIDE
};

typedef struct {
  int n;               /* grid points per axis, NX = NY = NZ */
  int interp_scheme;   /* 3 = INTERP_LAG3, 4 = INTERP_LAG4 */
  int precision;       /* GP_PREC_FP64 or GP_PREC_MIXED */
  int renorm_quat;     /* 0 = as the reference (no renormalisation) */
  double L;            /* LXMAX = LYMAX = LZMAX */
  double visc, rhof;   /* VISC, RHOF */
  double dtime;        /* DTIME */
  double ecw, muw;     /* wall restitution / friction (ECW, MUW) */
  double dpmax;        /* collision_initiation.f90:50 DPMAX; <= 0: 2*max(EMAJ) */
  int prefilter;       /* pairs proven disjoint skip the LM (their flag is
                          false in the reference too, so results do not
                          change): 0 none (LM on every pair); 1 centres
                          farther apart than the sum of the largest
                          semi-axes; 2 (recommended) 1 + the Perram-Wertheim
                          function above 1 + 1e-6 */
  int xcd_remap;       /* 1: each XCD takes a contiguous range of the
                          advance's, sort's and prefilter's workgroups (one
                          region of the field per XCD L2); 0: the hardware's
                          round-robin placement, all XCDs in one region that
                          the shared cache then holds -- the faster at every
                          configuration of 10^6 particles and more measured.
                          No effect on results */
  int keep_fluid;      /* 1: keep UFAP..VORTZAP of the last advance
                          (gp_get_fluid_at_particles) */
  int timing;          /* 1: accumulate per-stage HIP-event times (gp_stats) */
  int graph;           /* 1: gp_step replays a captured hipGraph of two full
                          steps (8 with overlap), no host synchronisation
                          inside a step; 2: the same steps launched directly
                          on the streams, without capture (faster than
                          replaying on this stack, the 8-step pipelined
                          graph costing ~60 us per step in node launches);
                          in both, buffers are sized from a first eager step
                          with headroom, and an overflow is reported as
                          GP_ERR_NCLOSE_OVERFLOW at the next gp_sync */
  int brick;           /* the advance visits particles grouped in bricks of
                          brick^3 collision cells (compact stencil sets per
                          workgroup, less field re-fetch); 0 or < 0: in the
                          reference's cell order (the default).  No effect on
                          results */
  int lm_seed;         /* initial guess of the contact solve: 0 the centre
                          line (reference, ellipsoid_contact_detection.f90:
                          53-62); 1 the seed network (gp_seed_net_*), with a
                          centre-line re-solve when the seeded solve ends
                          unconverged.  The solve decides every contact */
  int lm_waves;        /* wavefronts of the contact-solve launch (one pair per
                          lane, pairs dealt lane-major); 0: automatic from the
                          last known count of solves (one per SIMD while they
                          leave lanes idle, two or four per SIMD when each
                          lane has several) */
  int overlap;         /* 1: in graph-mode gp_step (and gp_collision_advance),
                          the next step's advance of the particles outside
                          the contact solves runs on a second stream beside
                          the solves (8 steps per captured graph), and each
                          step's contact pairs that were also the previous
                          step's are solved ahead on a third stream, beside
                          the step's sort and filters (while the solves are
                          latency-bound: fewer than the lanes of one
                          wavefront per SIMD; not with lm_seed = 1 or
                          collision statistics); 2: the same without solving
                          ahead; 0: steps run one after another.
                          No effect on results */
  int cell_rule;       /* collision cell size when dpmax <= 0.  0: the
                          reference rule DPMAX = 2 max(EMAJ), DXBOX = 2 DPMAX
                          (collision_initiation.f90:50-75), refused with
                          GP_ERR_CELL_RULE when a species' largest semi-axis
                          exceeds DXBOX / 2 (pairs would be lost); 1: DPMAX
                          = 2 x the largest semi-axis of any species (every
                          pair found; lists differ from the reference's when
                          an oblate semi-axis is the largest); 2: the
                          reference rule unchecked (the reference's lists,
                          lost pairs included) */
  int tile;            /* > 0: the advance visits particles grouped in tiles
                          of tile^3 field cells: one workgroup per tile
                          stages the field nodes its stencils reach in LDS
                          and writes the fluid at its particles (each node
                          read from HBM once per tile instead of once per
                          stencil), then the advance reads those values;
                          overrides brick.  Pays off where a tile holds a
                          wavefront's worth of particles (many ensembles on
                          one field).  The staged block, (tile+S-1)^2
                          (tile+3) nodes of 6 values (S = 4 lag3, 5 lag4),
                          must fit 64 KiB: tile <= 10 in the mixed form, 8
                          (lag3) or 7 (lag4) in FP64; else GP_ERR_ARG.
                          Results: identical for every tile size and launch
                          mode; against tile = 0 they differ by rounding
                          only (the advance runs in a kernel of its own,
                          whose fused multiply-adds the compiler chooses
                          separately).  0: off */
  int reserved[1];
} gp_params;

typedef struct {
  int64_t nclose;       /* candidate pairs of the last collision pass */
  int64_t nprefilter;   /* pairs that passed the bounding-sphere test */
  int64_t ncontact;     /* pairs with flag = .true. and NVREL < ZERO */
  int64_t ncollision;   /* pairs resolved (one collision per particle) */
  int64_t lm_nfev;      /* function evaluations summed over LM solves */
  int64_t nwall;        /* wall bounces in the last advance */
  int64_t greedy_rounds;
  /* stage times accumulated with HIP events on the context's stream when
   * gp_params.timing = 1 (0 otherwise): advance (gather + RK2 + walls
   * + wrap + binning), sort (scan, stable scatter, pair counts), contact
   * (prefilter + LM), greedy (one-collision pass + impulses) */
  double t_advance_ms, t_sort_ms, t_contact_ms, t_greedy_ms;
  int64_t ntimed;       /* steps included in the stage times */
  int64_t nlm;          /* pairs that reached the LM (after every prefilter) */
  int64_t seed_fallbacks; /* network-seeded solves redone from the centre line */
  int64_t lm_waves;     /* wavefronts of the last contact-solve launch issued
                           (a replayed graph keeps the width it was captured
                           with; gp_params.lm_waves = 0 chooses it) */
  double t_solve_ms;    /* with timing: the contact-solve kernel alone (and
                           the network's seeds), part of t_contact_ms */
  int64_t nahead;       /* of nlm, the pairs of the last collision pass that
                           were solved ahead (gp_params.overlap = 1) */
} gp_stats;

/* lifetime ----------------------------------------------------------- */
int gp_create(const gp_params *p, gp_ctx **out);
int gp_destroy(gp_ctx *ctx);
const char *gp_last_error_string(const gp_ctx *ctx);
int gp_sync(gp_ctx *ctx);

/* data in / out ------------------------------------------------------- */
/* UFLU..VORTICITY_Z after FLUIDCOMM (main.f90:255-259); copied once.   */
int gp_bind_field(gp_ctx *ctx, const double *field);
/* the same from device memory (same layout, on the context's device):
 * for a field supplier that keeps the fluid on the GPU (gp_dns_bind).
 * Re-binding keeps the context's buffers, so captured step graphs stay
 * valid. (No reference counterpart: its field lives in host arrays.)    */
int gp_bind_field_device(gp_ctx *ctx, const double *field);
/* the same from planar device arrays without ghosts: velocity and
 * vorticity double[3][n][n][n] each (x fastest; n must be the context's),
 * the vorticity multiplied by vort_scale; ghosts (x too) with index
 * period n.  One pass.                                                   */
int gp_bind_field_planar(gp_ctx *ctx, int n, const double *uvw, const double *vort,
                         double vort_scale);
/* per-species constants (PARAM_PHYS EMAJ_PART, APR_PART, RHOP, IPXX..)  */
int gp_set_species(gp_ctx *ctx, int nspec, const double *rows);
/* PART(1:np) -> device; species[i] in [0, nspec)                       */
int gp_set_particles(gp_ctx *ctx, int np, const double *state, const int *species);
/* ensembles: group[i] >= 0 (reference order) assigns each particle to an
 * independent ensemble sharing the context's field; particles of different
 * ensembles never form candidate pairs, and within an ensemble the pair
 * list keeps the reference order (no counterpart in the reference, which
 * runs one ensemble per process; gp_set_particles resets to one ensemble) */
int gp_set_groups(gp_ctx *ctx, int np, const int *group);
/* device -> PART(1:np) in the order given to gp_set_particles           */
int gp_get_particles(gp_ctx *ctx, int np, double *state);
/* fluid velocity and vorticity at x^n of the last advance (UFAP..VORTZAP),
 * double[6][np]; requires gp_params.keep_fluid = 1 at create time       */
int gp_get_fluid_at_particles(gp_ctx *ctx, int np, double *fl);

/* hot-path entry points ------------------------------------------------ */
/* INTERPH (interpH.f90:26-145) of field component comp (0..5) at nint
 * host positions; scheme 3 or 4.                                       */
int gp_interph(gp_ctx *ctx, int scheme, int nint, const double *x, const double *y,
               const double *z, int comp, double *out);
/* gp_interph without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid it).  The points and results are
 * staged in a fixed device array of GP_INTERPH_NOALLOC_MAX points (8 MiB
 * per GPU, held until the process ends); more points are refused with
 * GP_ERR_ARG.  All contexts on one GPU share that array: calls that
 * overlap on one GPU (from several host threads) race.                  */
#define GP_INTERPH_NOALLOC_MAX 262144
int gp_interph_noalloc(gp_ctx *ctx, int scheme, int nint, const double *x, const double *y,
                       const double *z, int comp, double *out);
/* INTERPH with the field passed in, as the reference passes UAIN: one
 * component UAIN(1:n,-1:n+2,-1:n+2) of the context's grid, FP64 gather.
 * (Copies the component to the device on every call; the step itself
 * uses the field bound by gp_bind_field.)                              */
int gp_interph_field(gp_ctx *ctx, int scheme, const double *uain, int nint, const double *x,
                     const double *y, const double *z, double *out);
/* ADV_PARTICLE_ELLP (adv_particle_new.f90:7-704, called at main.f90:466)
 * on the device-resident particles, including BOUNDARY_PARTICLE.       */
int gp_adv_particle_ellp(gp_ctx *ctx, int ncycle);
/* COLLISION_DETECTION_CONNECTIVITY (collision_detection_connectivity.f90:
 * 7-232) on the device-resident positions: 1-based HOME/NEAR in the
 * reference order.  Returns GP_ERR_NCLOSE_OVERFLOW (with *nclose set) if
 * the list does not fit.                                                */
int gp_collision_detection(gp_ctx *ctx, int64_t nclosemax, int *home, int *near_,
                           int64_t *nclose);
/* ellipsoid_contact_detection (ellipsoid_contact_detection.f90:1-155) on
 * npairs (1-based) pairs of device-resident particles, with the minimum
 * image of collision_interaction.f90:221-237.  Any output may be NULL.
 * With gp_params.lm_seed = 1 and a network set, each pair is solved as
 * the step solves it (network seed through the gate of gp_seed_net_gate,
 * nfev counting both attempts).  EPT1, EPT2 and cnormal are those of the
 * solve's final iterate whatever the flag.                               */
int gp_contact_batch(gp_ctx *ctx, int64_t npairs, const int *home, const int *near_,
                     unsigned char *flag, double *ept1, double *ept2, double *cnormal,
                     int *nfev);
/* The approaching contacts (flag true and NVREL < ZERO) of the last
 * collision pass, in the order COLLISION_INTERACTION's loop meets them
 * (collision_interaction.f90:194-419): 1-based HOME and NEAR, and
 * resolved = 1 where the one-collision pass resolved the pair (both
 * particles still free when the loop reached it, :207, :348, :397-398).
 * The reference prints these pairs (print_colliding_pair.f90) rather than
 * returning them.  *n is the number of contacts, also when it exceeds
 * nmax (then nmax are written).  Valid after gp_collision or an eager
 * gp_step and until the next advance; GP_ERR_STATE otherwise.            */
int gp_get_contacts(gp_ctx *ctx, int64_t nmax, int *home, int *near_, unsigned char *resolved,
                    int64_t *n);
/* COLLISION (collision.f90:7-201, called at main.f90:483) at NPROC = 1. */
int gp_collision(gp_ctx *ctx, int ncycle, double time);
/* nsteps x (ADV_PARTICLE_ELLP + COLLISION), asynchronous on the context's
 * stream; errors surface at the next gp_sync / gp_get_particles.        */
int gp_step(gp_ctx *ctx, int nsteps);
/* COLLISION (collision.f90:7, called at main.f90:483) of cycle ncycle
 * followed by ADV_PARTICLE_ELLP (adv_particle_new.f90:7, main.f90:466) of
 * cycle ncycle_next, in one asynchronous call: the advance of every
 * particle outside this pass's contact solves runs on a second stream
 * beside the solves, the solves' particles are advanced after the
 * one-collision pass.  Same results as gp_collision then
 * gp_adv_particle_ellp.  For a caller stepping once per call (the Fortran
 * drop-in's resident mode).  The first call after gp_set_particles sizes
 * the pair lists with headroom and synchronises; later calls do not (an
 * overflow surfaces as GP_ERR_NCLOSE_OVERFLOW at the next gp_sync).  With
 * collision statistics or stage timers on, the two run one after the
 * other.                                                                 */
int gp_collision_advance(gp_ctx *ctx, int ncycle, double time, int ncycle_next);
/* DTIME for the following steps (CHECK_CFL changes it every cycle when the
 * fluid evolves, check_cfl.f90:92); a captured step graph is re-captured */
int gp_set_dtime(gp_ctx *ctx, double dtime);
int gp_get_stats(gp_ctx *ctx, gp_stats *out);

/* STAT_PARTICLE, level 1 (stat_particle.f90:169-450) over the resident
 * particles: out[s][k-1] = MEAN_PART(k, J, IDP), k = 1..41, species rows
 * class-major with npoly size indices per class (s = (J-1) npoly + IDP-1,
 * nspec = the species count, a multiple of npoly).  ref_norm = 1 divides as
 * the reference does, every class by the last class's counts per size
 * index (stat_particle.f90:436-449: NPM(IDP) keeps the last class loop's
 * values); ref_norm = 0 divides each row by its own count.  The sums run in
 * another order than the reference's particle loop.                      */
int gp_particle_stats(gp_ctx *ctx, int nspec, int npoly, int ref_norm, double *out);
/* STAT_PARTFLUID (stat_partfluid.f90:113-381) over the resident particles
 * and the fluid at them of the last advance (needs gp_params.keep_fluid =
 * 1 and an advance since gp_set_particles): out[s][k-1] =
 * MEAN_PARTFLUID(k, J, IDP), k = 1..36 (Rep, CD, CL, drag and lift in the
 * lab frame and their norms, Re_R, pitching / rotation torque
 * coefficients, the torques in the body frame and their norms, TAUPF,
 * GRAVITY, incidence angle, major axis, 1/tau_p, particle velocity and
 * body angular velocity, count), with the Zastawny fits that routine fixes;
 * species rows, npoly and ref_norm as for gp_particle_stats (the routine
 * divides by the same last-class counts, :362-381).  PMASS, dequiv and
 * TAUPF come from the rows' RHOP, EMAJ, APR columns as the routine forms
 * them.                                                                  */
int gp_partfluid_stats(gp_ctx *ctx, int nspec, int npoly, int ref_norm, double *out);
/* STAT_PARTICLE's PDFs (LEVEL1_STPDF, stat_particle.f90:536-645, on the
 * cycles that are multiples of 500) over the resident particles:
 * out[s][k][b] = MEAN_PART_PDF(k+1, J, IDP, b+1) (counts added to the
 * routine's ZERO = 1e-33 start, :164), k = 0..11: u, v,
 * w in [-2, 2]; body-frame spin x, y, z in [-2000, 2000]; lab-frame spin
 * R(q) omega in [-2000, 2000]; the body x axis in the lab frame R(q) e_x
 * in [-1, 1]; GP_NPDF bins, bin = int((value - min)/width) + 1 clamped to
 * [1, GP_NPDF] as the routine bins them.  Species rows as for
 * gp_particle_stats.  The caller adds these to MEAN_TIME_PART_PDF.       */
#define GP_NPDF 100
int gp_particle_pdf(gp_ctx *ctx, int nspec, double *out);
/* STAT_PARTICLE level 2 (LEVEL2_STPAR, stat_particle.f90:737-826), the
 * Lagrangian velocity autocorrelation.  gp_lagrangian_origin, at cycle
 * NT0(NLGR) * FOUT2, keeps for correlation slot NLGR - 1 (0 <= slot <
 * GP_NLGR) every particle's u - MEAN_PART(1:3, J, IDP) (UPT0..WPT0, :760-
 * 770), mean[s][3] given per species row (MEAN_PART(1:3) of the same
 * cycle's level 1: gp_particle_stats with ref_norm = 1, entries 1-3).
 * gp_lagrangian_corr, on each later statistics cycle, returns per row
 * out[s][3] = the sums over the row's particles of (u - mean) * u_t0 /
 * npm[s] (:790-805), the increments of RPX/RPY/RPZ_LOC(NTLGR, J, NLGR,
 * IDP), with npm[s] = NPM(IDP) as the routine holds it (the last class's
 * count for size IDP).  Sums in another order than the particle loop.    */
#define GP_NLGR 16
int gp_lagrangian_origin(gp_ctx *ctx, int slot, int nspec, const double *mean);
int gp_lagrangian_corr(gp_ctx *ctx, int slot, int nspec, const double *mean, const double *npm,
                       double *out);
/* zero the accumulated stage times                                       */
int gp_reset_timers(gp_ctx *ctx);
/* introspection for benchmarks: collision grid of the context           */
int gp_collision_grid(gp_ctx *ctx, int *nbox_xyz, double *box_min_size);
/* collisions each particle took part in since gp_set_particles (or the
 * last reset), in reference order (the per-particle collision statistics
 * of collision_interaction.f90:330-345); reset != 0 zeroes them */
int gp_get_collision_counts(gp_ctx *ctx, int np, int *counts, int reset);

/* Level-1 collision statistics of COLLISION_INTERACTION (MEAN_PART_COL,
 * collision_interaction.f90:330-345, 409, 423-441; with LEVEL1_STPAR and
 * STAT_TIME, :453).  fout2 > 0 turns them on (fout2 <= 0 off): every
 * collision pass whose cycle number is a multiple of fout2 sums, over the
 * contacts the ordered pass reaches with both particles free, per class J
 * and size pair (IDP1 of HOME, IDP2 of NEAR) with species rows class-major
 * (s = (J-1) npoly + IDP-1; rows beyond nclass * npoly are not counted):
 *   1 contacts, 2 collisions (NVREL < 0), 3 2 x (2) / DTIME,
 *   4 sum of |EPT1 - EPT2|, 5 (4) / (1), 6 sum of NVREL, 7 sum of |VREL|,
 *   8 sum of acos(NVREL / |VREL|),
 * every sum starting from the reference's ZERO = 1e-33 (so a bin with no
 * contact reads 1e-33 in 1, 2, 4, 6-8 and 1 in 5, as the reference's).
 * gp_collision(ncycle) tests its own ncycle; gp_step numbers its passes
 * on from the last one (the first is ncycle here).  Turning them on (or
 * off) clears the sums and re-captures the step graph.  The sums run in
 * another order than the reference's list loop.                         */
#define GP_CSTAT_MAX_BINS 65536
int gp_set_collision_stats(gp_ctx *ctx, int nclass, int npoly, int fout2, int ncycle);
/* last[((J-1) npoly + IDP1-1) npoly + IDP2-1][k-1] = MEAN_PART_COL(k,
 * IDP1, IDP2) of class J at the last statistics pass; time_sum[...] the
 * sum over statistics passes from ZERO (MEAN_TIME_PART_COL,
 * collision_initiation.f90:172; kept per class: the reference adds all
 * classes into one array); nstat the number of
 * statistics passes.  Any output may be NULL.                           */
int gp_get_collision_stats(gp_ctx *ctx, double *last, double *time_sum, int64_t *nstat);

/* contact-seed network -------------------------------------------------
 * No reference counterpart: a 12 -> 64 -> 64 -> 4 ReLU network (bf16
 * weights, fp32 accumulate on the matrix cores).  From a candidate pair's
 * relative pose and semi-axes (body frame of the first particle) it
 * predicts the common normal the contact solve converges to when it finds
 * a contact (outputs 0..2) and the logit of finding one (output 3).  It
 * only seeds the solve (gp_params.lm_seed = 1), and only where the
 * predicted contact probability exceeds 0.9.  Parameters travel as one
 * float vector: W1 (64x12), b1 (64), W2 (64x64), b2 (64), W3 (4x64),
 * b3 (4), row-major.  Feature rows: R1^T (b2 - b1)/s, a1/s, a2/s, rotation
 * vector of q1* q2, s = max(a1) + max(a2).  Target rows: the converged
 * unit normal in body frame 1 and the contact flag (0 / 1); loss per row
 * flag |y0..2 - t0..2|^2 + softplus(y3) - flag y3.                       */
#define GP_SEED_NET_NPARAM 5252

typedef struct {
  int64_t npairs;
  double loss;           /* mean row loss over the last epoch */
  double label_ms;       /* solver labelling of the training pairs */
  double train_ms;       /* forward + backward + Adam on the GPU */
  double nfev_ref_mean;  /* evaluations per solve from the centre line */
  double nfev_net_mean;  /* ... from the network seed (fallbacks counted) */
  int64_t nfev_ref_max, nfev_net_max;
  int64_t flags_ref, flags_net, flag_agree;   /* contact flags per seed */
  int64_t fallbacks;     /* seeded solves redone from the centre line */
  int64_t no_seed;       /* pairs where the network gave no usable normal */
  double cn_min_dot;     /* min |cn_ref . cn_net| over pairs flagged by both */
  double reserved[4];
} gp_seed_report;

int gp_seed_net_set(gp_ctx *ctx, const float *theta, int nparam);
int gp_seed_net_get(gp_ctx *ctx, float *theta, int nparam);
/* y[n][4] = network(x[n][12]) */
int gp_seed_net_forward(gp_ctx *ctx, int64_t n, const float *x, float *y);
/* gradient (same layout as theta) of the mean row loss; t[n][4] */
int gp_seed_net_loss_grad(gp_ctx *ctx, int64_t n, const float *x, const float *t,
                          float *grad, double *loss);
/* Adam (lr, beta 0.9 / 0.999, eps 1e-8) over caller rows (t[n][4]),
 * batches in order */
int gp_seed_net_train(gp_ctx *ctx, int64_t n, const float *x, const float *t, int epochs,
                      int batch, float lr, double *loss);
/* npairs synthetic tri-axial pairs (semi-axis ratios U[1,4], uniform
 * orientations, centre distance U[0.2,1] x sum of largest semi-axes, not
 * proven disjoint) labelled by the FP64 solve from the centre line; x, t,
 * nfev, poses (n x 20: b1, q1, a1, b2, q2, a2) may be NULL */
int gp_seed_net_pairs(gp_ctx *ctx, int64_t n, uint64_t seed, float *x, float *t, int *nfev,
                      double *poses);
/* label npairs such pairs on the GPU and train on them */
int gp_seed_net_fit(gp_ctx *ctx, int64_t npairs, uint64_t seed, int epochs, int batch,
                    float lr, gp_seed_report *rep);
/* solve fresh pairs from both seeds: evaluation counts, flag agreement */
int gp_seed_net_eval(gp_ctx *ctx, int64_t npairs, uint64_t seed, gp_seed_report *rep);
/* Acceptance gate of a network-seeded solve (no reference counterpart):
 * the seeded attempt runs at most maxfev evaluations (1..300) and is kept
 * only if it converged with its final normal within keep_angle radians of
 * the seed's; otherwise the pair is solved from the centre line as the
 * reference does (ellipsoid_contact_detection.f90:77-87, maxfev 300).
 * Defaults: 6 evaluations, 0.02 rad.                                    */
int gp_seed_net_gate(gp_ctx *ctx, int maxfev, double keep_angle);

/* on-disk formats ------------------------------------------------------
 * Host only (no context, no GPU): the files either side of the step, so
 * that the path reads and writes what existing runs hold.
 *
 * Per-class particle file, ISAVEPART = 2: written by SAVE_PARTICLE as
 * part_pJJ.end / part_pJJ_tNNNNNNNN.bin (save_particle.f90:116-185), read
 * back as part_pJJ.ini by INITIATION_PARTICLE_POSITION
 * (initiation_particle_position.f90:430-514).  Fortran unformatted
 * sequential records (4-byte length before and after each): NPART_FULL;
 * NPCLASS(J,1:npoly); RHOP(J,1:npoly); EMAJ_PART(J,1:npoly); APR_PART(J);
 * then 14 records of npart doubles.  cols is double[14][npart]: the state
 * components (layout above) followed by the size index IDP as a double.
 * Records over 2^31 - 1 bytes (npart > 268,435,455) are refused
 * (GP_ERR_ARG).                                                          */
int gp_part_file_write(const char *path, int64_t npart, int npoly, const int *npclass,
                       const double *rhop, const double *emaj, double apr, const double *cols);
/* npart and npoly (length of the NPCLASS record) of a particle file      */
int gp_part_file_info(const char *path, int64_t *npart, int *npoly);
/* reads a file whose npart and npoly match (GP_ERR_IO otherwise)        */
int gp_part_file_read(const char *path, int64_t npart, int npoly, int *npclass, double *rhop,
                      double *emaj, double *apr, double *cols);
/* Velocity-component file of SAVE_MPIIO / READ_MPIIO (uf.end, vf.end,
 * wf.end, uf_tNNNNNNNN.bin ...; MPI_structures.f90:434-493, 571-621, names
 * at save_fluid.f90:141-186 and initiation_fluid.f90:275-300): three int32
 * NX NX NX, then NX^3 doubles in Fortran order (x fastest, no ghosts).   */
int gp_field_file_write(const char *path, int n, const double *f);
int gp_field_file_read(const char *path, int n, double *f);

/* spectral DNS field supplier ---------------------------------------------
 * The step before the path when the fluid evolves (SOLVE_FLUID = 1): the
 * reference's pseudo-spectral Navier-Stokes step ADV_FLUID
 * (adv_fluid.f90:78-234: forward transform of UFLU, two-stage
 * integrating-factor Runge-Kutta on the rotational-form nonlinear term
 * of NONLINEAR_TERMS (nonlinear_terms.f90:96-206), PROJ_DIVFREE
 * (proj_divfree.f90:46-60), the smooth dealiasing FILTER of MESHING
 * (meshing.f90:151-180), backward transform), with the STEADY forcing of
 * ADV_FORCING / ADD_FORCING3 (adv_forcing.f90:33-125, add_forcing3.f90:
 * 40-58; Ornstein-Uhlenbeck amplitudes on KFORCE_MIN <= |k| <= KFORCE_MAX
 * from the reference's RAN1 / GAUSS sequence, initial amplitudes 0 as
 * INITIATION_FORCING_NEW sets them for INIT_FLUID_VELOCITY <= 2) and
 * CHECK_CFL (check_cfl.f90:61-110).  Velocity and spectral state stay in
 * device memory (rocFFT, double precision); gp_dns_bind hands velocity and
 * a spectrally computed vorticity to a particle context without a host
 * copy.  Grids: n^3 points, the transform grid x = (i-1) L / n; arrays
 * double[3][n][n][n], x fastest.                                         */
typedef struct gp_dns gp_dns;
typedef struct {
  int n;                 /* NX = NY = NZ (even) */
  double L;              /* LXMAX = LYMAX = LZMAX */
  double visc, dtime;    /* VISC, DTIME */
  int forcing;           /* 1: STEADY forcing */
  int seed;              /* IDFORCE (RAN1 seed; <= 0 re-initialises the shuffle) */
  double kforce_min, kforce_max, sigma_force, time_force;
  int reuse_spectral;    /* 1: start each step from the spectral state the
                            previous one ended with instead of transforming
                            UFLU again (adv_fluid.f90:86-94 does; the two
                            differ by rounding only, the filter having
                            removed the modes a round trip could change):
                            one transform of seven less per step */
  int reserved[3];
} gp_dns_params;
int gp_dns_create(const gp_dns_params *p, gp_dns **out);
int gp_dns_destroy(gp_dns *d);
const char *gp_dns_error_string(const gp_dns *d);
/* UFLU, VFLU, WFLU (double[3][n][n][n]) -> device                         */
int gp_dns_set_velocity(gp_dns *d, const double *uvw);
int gp_dns_get_velocity(gp_dns *d, double *uvw);
/* curl of the spectral velocity, i k x U, back on the grid               */
int gp_dns_get_vorticity(gp_dns *d, double *w);
/* nsteps x ADV_FLUID at the current DTIME                                 */
int gp_dns_step(gp_dns *d, int nsteps);
/* CHECK_CFL with NPROC = 1: DTIME = min(Courant 0.05, Fourier 1.0,
 * dtime_user) with DX = L/(n-1) as the reference's mesh; sets and returns it */
int gp_dns_cfl(gp_dns *d, double dtime_user, double *dtime);
/* EPS_FLU of the last step's first stage (nonlinear_terms.f90:124-128),
 * the velocity's largest magnitudes per component, and the forcing mode
 * count                                                                  */
int gp_dns_info(gp_dns *d, double *eps, double *umax3, int *nforce);
/* velocity and vorticity -> the particle context's field (its n must be
 * the DNS's), device to device.  mesh_vorticity = 1 scales the vorticity
 * by (n-1)/n: the curl of the velocity as the particles see it on the
 * reference's mesh XMESH = (i-1) L/(n-1); 0 binds the transform grid's.  */
int gp_dns_bind(gp_dns *d, gp_ctx *ctx, int mesh_vorticity);

#ifdef __cplusplus
}
#endif
#endif
