/*
 * gazpart_oracle.h -- CPU FP64 restatement of GAZPART's Lagrangian ellipsoid
 * step (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the HIP path: only tests/, the smoke entry
 * and the CPU-baseline leg of the benchmark may load it.  It is never linked
 * into, or called by, the product library.
 *
 * Every routine follows one reference subroutine (paths relative to the
 * reference's source/ directory) and keeps its arithmetic order, so that on
 * identical inputs it reproduces the Fortran FP64 result (the reference is
 * compiled with -fdefault-real-8 -fdefault-double-8).
 *
 * Layouts
 *   field  : 6 consecutive components (u, v, w, wx, wy, wz), each in the
 *            reference array shape UAIN(1:n, -1:n+2, -1:n+2), x fastest
 *            (allocate_arrays.f90:108-118, NGHTCELL = 2).
 *   state  : component-major SoA, double[OC_NSTATE][np] with components
 *            x y z u v w qa qb qc qd ox oy oz (omega in the body frame).
 *   species: double[nspec][OC_SP_STRIDE], one row per (class, size) pair,
 *            built by oc_species_spheroid / oc_species_triaxial.
 */
#ifndef GAZPART_ORACLE_H
#define GAZPART_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_X = 0, OC_Y, OC_Z, OC_U, OC_V, OC_W, OC_QA, OC_QB, OC_QC, OC_QD,
       OC_OX, OC_OY, OC_OZ, OC_NSTATE };

/* species row layout */
enum { SP_A = 0, SP_B, SP_C, SP_PMASS, SP_DEQ, SP_IXX, SP_IYY, SP_IZZ,
       SP_GRAV, SP_ECP, SP_MUP, SP_RHOP, SP_EMAJ, SP_APR, SP_STRIDE };
#define OC_SP_STRIDE SP_STRIDE

/* status codes (shared numbering with the HIP library) */
enum {
  OC_OK = 0,
  OC_ERR_NAN_DRAG = 1,      /* fluid_force_ellipsoid.f90:150-158 */
  OC_ERR_NAN_OMEGA = 2,     /* fluid_torque_ellipsoid.f90:178-188 */
  OC_ERR_KE_COLLISION = 3,  /* collision_response.f90:408-436 */
  OC_ERR_KE_WALL = 4,       /* wall_ellipsoid_rebound.f90:372-384 */
  OC_ERR_OUT_OF_GRID = 5,   /* collision_detection_connectivity.f90:81-89 */
  OC_ERR_NCLOSE_OVERFLOW = 6,
  OC_ERR_ARG = 10
};

typedef struct {
  int n;              /* grid points per axis (NX = NY = NZ) */
  int interp_scheme;  /* 3 = INTERP_LAG3, 4 = INTERP_LAG4 */
  int renorm_quat;    /* 0 = reference (no renormalisation) */
  int pad_;
  double L;           /* LXMAX = LYMAX = LZMAX */
  double dx;          /* L / (n - 1), meshing.f90:43 */
  double visc, rhof, dtime;
  double ecw, muw;    /* wall restitution / friction */
} oc_params;

typedef struct {
  int nxbox, nybox, nzbox, nbox;
  double xminbox, yminbox, zminbox;
  double dxbox, dybox, dzbox;
} oc_cgrid;

/* ---- species constants (reference formulas) ---------------------------- */
void oc_species_spheroid(double emaj, double apr, double rhop, double grav,
                         double ecp, double mup, double *row);
void oc_species_triaxial(double a, double b, double c, double rhop,
                         double grav, double ecp, double mup, double *row);

/* ---- field / interpolation -------------------------------------------- */
int64_t oc_field_size(int n);   /* n*(n+4)*(n+4) values per component */
void oc_fill_ghosts(int n, double *comp);  /* y/z ghosts, index period n */
int oc_interph(int scheme, int n, double dx, const double *uain, int nint,
               const double *xint, const double *yint, const double *zint,
               double *uint_out);

/* ---- particle kernels --------------------------------------------------- */
void oc_transform_basis_vec(const double q[4], double v[3]);
void oc_quaternion_integration(double q[4], double ox, double oy, double oz,
                               double dt);
int oc_fluid_force(const oc_params *p, const double *sp, double up, double vp,
                   double wp, double uf, double vf, double wf,
                   const double q[4], double fdrag[3], double flift[3]);
int oc_fluid_torque(const oc_params *p, const double *sp, double ufap,
                    double vfap, double wfap, double upn, double vpn,
                    double wpn, double vxap, double vyap, double vzap,
                    double ox, double oy, double oz, const double q[4],
                    double tpitch[3], double trot[3]);
void oc_wall_test(const oc_params *p, const double *sp, double x, double y,
                  double z, const double q[4], double wall_normal[3],
                  double impact_arm[3], double *delta_x, int *intersect);
int oc_wall_rebound(const oc_params *p, const double *sp, double *u,
                    double *v, double *w, double *ox, double *oy, double *oz,
                    const double q[4], const double arm[3],
                    const double nrm[3]);
int oc_adv_particle_ellp(const oc_params *p, const double *field,
                         const double *spec, int np, double *state,
                         const int *species, double *fluid_at_p,
                         int nthreads);
void oc_boundary_particle(const oc_params *p, int np, double *state);

/* ---- collision ---------------------------------------------------------- */
void oc_collision_grid_init(const oc_params *p, double dpmax, oc_cgrid *g);
void oc_box_connectivity(const oc_cgrid *g, int periodic, int *con,
                         int *ncon);
int oc_collision_detection(const oc_cgrid *g, const int *con, const int *ncon,
                           int np, const double *x, const double *y,
                           const double *z, int64_t nclosemax, int *home,
                           int *near_, int64_t *nclose);
int oc_contact_detection(const double *spi, const double *spj,
                         const double bi[3], const double qi[4],
                         const double bj[3], const double qj[4],
                         const double *seed, int *flag, double ept1[3],
                         double ept2[3], double cnormal[3], int *nfev,
                         double xsol[2]);
int oc_collision_response(const double *sp1, double *s1, double *s2,
                          const double q1[4], const double q2[4],
                          const double r1[3], const double r2[3],
                          const double nrm[3]);
int oc_collision_interaction(const oc_params *p, const double *spec, int np,
                             double *state, const int *species,
                             int64_t nclose, const int *home,
                             const int *near_, int *flag_out,
                             int *collided_out, int *ncoll, int nthreads);
/* the same pass with the level-1 collision statistics of
 * collision_interaction.f90:330-345, 409, 423-441: col (nbin x 8, bin =
 * species[HOME] * npoly + species[NEAR] % npoly) receives MEAN_PART_COL(1:8)
 * of the pass, every entry starting from the reference's ZERO = 1e-33 */
int oc_collision_interaction_stats(const oc_params *p, const double *spec, int np,
                                   double *state, const int *species,
                                   int64_t nclose, const int *home,
                                   const int *near_, int *flag_out,
                                   int *collided_out, int *ncoll, int nthreads,
                                   int npoly, int nbin, double *col);
int oc_step(const oc_params *p, const double *field, const double *spec,
            int np, double *state, const int *species, const oc_cgrid *g,
            const int *con, const int *ncon, int64_t nclosemax, int *home,
            int *near_, int64_t *nclose, int *ncoll, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
