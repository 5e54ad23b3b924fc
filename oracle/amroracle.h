/*
 * amroracle.h -- CPU restatement (plain C, FP64) of the reference's two-phase
 * AMR-LBM time step. TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
 * library is compared against, never the thing measured or shipped. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Array layouts follow the reference exactly (Fortran column-major):
 *   h, g : (9, 0:nx+1, 0:ny+1)   population index fastest
 *   C    : (-1:nx+2, -1:ny+2)
 *   P,Ux,Uy : (nx, ny)
 * Block ids are 1-based as in Blk(MaxBlks) (Parameters.f90:39).
 */
#ifndef AMRORACLE_H
#define AMRORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny;          /* cells per block side (Parameters.f90:12-13) */
    int L0;              /* root block extent in finest cells (Parameters.f90:15) */
    int nroots;          /* root blocks stacked in y (AMR_Rising_2D_OMP.f90:54-57) */
    int min_levels;      /* Parameters.f90:17 */
    int max_levels;      /* Parameters.f90:16 */
    int max_blocks;      /* Parameters.f90:19 */
    double M0;           /* D2Q9_Mod.f90:10 */
    double W, sigma, beta, kappa, taul, tauh, rhol, rhoh, gy;  /* D2Q9_Mod.f90:13-16 */
} orc_params;

/* neighbour pointer: the reference's TYPE(Ptr) (Parameters.f90:23-25).
 * kind 0: points at a fixed integer (root-block storage or a wall code);
 * kind 1: aliases Blk(a)%child(b) (Neighbors.f90:30,35). */
typedef struct { int kind, a, b; } orc_ptr;

typedef struct {
    int level, dx, I, J, parent;
    int child[4];
    int has_child;               /* child(:) allocated */
    orc_ptr nghbr[8];
    int isleaf, hasleaf, refin, deref;
    double *h, *g, *C, *P, *Ux, *Uy;   /* NULL when not allocated */
} orc_block;

typedef struct {
    orc_params p;
    orc_block *blk;              /* blk[1..max_blocks] */
    int nblks, nleaf, *leaf;     /* LeafBlks(1:nLeaf) */
    int levelmin, levelmax, kk, T;
    int *freeq, fhead, ftail;    /* FIFO of freed child groups (Parameters.f90:41-45) */
    int overflow;                /* MaxBlks exceeded (AMR_Rising_2D_OMP.f90:527-531) */
    int level_errors;            /* unexpected level jumps met by the ghost fills */
} orc_mesh;

/* parameter derivation of Set_Rising_Bubble_Parameters (AMR_Rising_2D_OMP.f90:235-263) */
void orc_reference_physics(orc_params *p, double D, int Bo, double Ar, double gy, double W,
                           double rhol, double rhoh, double visc_ratio);

orc_mesh *orc_create(const orc_params *p);
void orc_destroy(orc_mesh *m);

/* set-up: roots, Initial_Mesh, analytic bubble, init-regrid loop
 * (AMR_Rising_2D_OMP.f90:51-67, 79-231, 493-584) */
int orc_setup_rising_bubble(orc_mesh *m, double R, double X0, double Y0);

/* analytic phase field of the set-up centred at (X0,Y0) on every leaf, both
 * ghost rings included (AMR_Rising_2D_OMP.f90:174-185) */
void orc_set_bubble_c(orc_mesh *m, double R, double X0, double Y0);

/* the phases of one step (AMR_Rising_2D_OMP.f90:302-364) */
void orc_determine_min_max_levels(orc_mesh *m);
void orc_fill_ghost_cells_c2(orc_mesh *m);
void orc_collide_all(orc_mesh *m);
void orc_fill_ghost_cells_hg(orc_mesh *m);
void orc_stream_all(orc_mesh *m);
void orc_macroscopic_h_all(orc_mesh *m);
void orc_fill_ghost_cells_c(orc_mesh *m);
void orc_macroscopic_g_all(orc_mesh *m);
void orc_refine_derefine_mesh(orc_mesh *m, double refin_tol, double deref_tol);
/* whole step; regrid after step T when (T+1) % regrid_every == 0 */
void orc_step(orc_mesh *m, int regrid_every);
void orc_run(orc_mesh *m, int nsteps, int regrid_every);
void orc_set_step(orc_mesh *m, int T);   /* the step counter T written to snapshots */

/* FP64 diagnostics over leaf interiors: rho_min, rho_max, max|Uy|, mass
 * (AMR_Rising_2D_OMP.f90:464-483) and bubble area, centroid x, y, rise velocity
 * (weights (1-C) dx^2), non-finite count */
void orc_diagnostics(const orc_mesh *m, double out[9]);

/* per-block kernels (LBM_Subroutines.f90), reference layouts */
void orc_collision_block(const orc_params *p, int dx, const double *C, const double *P,
                         const double *Ux, const double *Uy, double *h, double *g);
void orc_macro_h_block(const orc_params *p, const double *h, double *C);
void orc_macro_g_block(const orc_params *p, int dx, const double *g, const double *C,
                       double *P, double *Ux, double *Uy);
void orc_perfect_shift_block(const orc_params *p, double *f);
void orc_lax_wendroff_block(const orc_params *p, double cfl, double *f);
double orc_err_interface(const orc_params *p, int dx, const double *C);

/* snapshot exchange in the reference harness's binary layout */
int orc_load_snapshot(orc_mesh *m, const char *path);
int orc_save_snapshot(const orc_mesh *m, const char *path);
int orc_nghbr_value(const orc_mesh *m, int n, int k);   /* k = 1..8 */

/* interpolation primitives (Interpolations.f90), exposed for closed-form tests */
void orc_biquadratic_from_fine(int i2, int j2, int X2, int Y2, const double *in, double *out);
void orc_biquadratic_from_coarse(int I2, int J2, int x2, int y2, const double *in, double *out);
void orc_bilinear_from_fine(int i2, int j2, int X2, int Y2, const double *in, double *out);
void orc_bilinear_interp_extrap(int nx, int ny, const double *in, double *out);
void orc_mrt_collision(double s9, const double geq[9], double g[9]);

#ifdef __cplusplus
}
#endif
#endif
