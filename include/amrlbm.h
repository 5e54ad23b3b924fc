/*
 * amrlbm.h -- C-ABI of the MI355X (gfx950) two-phase AMR-LBM step.
 *
 * The library owns the mesh on one GPU: block table, neighbour table and the
 * per-block arrays of the reference (h, g, C, P, Ux, Uy), laid out block-major
 * structure-of-arrays in HBM. Each entry point below replaces one subroutine of
 * AbbasHub/AMR-LBM-OpenMP-2D (cited beside it) with the same meaning; the
 * Fortran shim in amr-lbm-openmp-2d_amd/fortran binds them with ISO_C_BINDING
 * under the reference's own subroutine names.
 *
 * Conventions
 *  - every function returns 0 (AMRLBM_OK) or an AMRLBM_ERR_* code; the text of
 *    the last error of the calling thread is amrlbm_last_error();
 *  - block ids are the reference's 1-based Blk(n) ids (Parameters.f90:39);
 *  - host arrays use the reference's layouts (column-major, population index
 *    fastest): h,g (9,0:nx+1,0:ny+1); C (-1:nx+2,-1:ny+2); P,Ux,Uy (nx,ny);
 *  - a mesh is not re-entrant: one thread at a time per mesh. All device work
 *    of a mesh is ordered on one HIP stream of default priority; entry points
 *    that return data synchronise that stream.
 */
#ifndef AMRLBM_H
#define AMRLBM_H
#ifdef __cplusplus
extern "C" {
#endif

#define AMRLBM_OK 0
#define AMRLBM_ERR_HIP 1        /* a HIP runtime call failed */
#define AMRLBM_ERR_ARG 2        /* invalid argument */
#define AMRLBM_ERR_CAPACITY 3   /* MaxBlks exceeded (AMR_Rising_2D_OMP.f90:527-531 stops the run) */
#define AMRLBM_ERR_TOPOLOGY 4   /* 2:1 balance violated / unresolvable neighbour (Fill_Left_hg:101-103 prints) */
#define AMRLBM_ERR_NONFINITE 5  /* non-finite phase field found by amrlbm_output */
#define AMRLBM_ERR_STATE 6      /* call out of order (e.g. no mesh set up) */

typedef struct amrlbm_mesh amrlbm_mesh;

/* Compile-time PARAMETERs and module variables of the reference
 * (Parameters.f90:12-19, D2Q9_Mod.f90:10-16), passed at run time. */
typedef struct {
    int nx, ny;          /* cells per block side; this build supports 16 x 16 */
    int L0;              /* root block extent in finest cells */
    int nroots;          /* root blocks stacked in y (x is periodic) */
    int min_levels, max_levels;
    int max_blocks;      /* MaxBlks */
    double M0, W, sigma, beta, kappa, taul, tauh, rhol, rhoh, gy;
} amrlbm_params;

/* ---- life cycle ---------------------------------------------------------- */
int amrlbm_create(const amrlbm_params *p, amrlbm_mesh **out);   /* module parameters (Parameters.f90) */
/* The FP32 production form of the same mesh (no counterpart in the reference,
 * which is FP64 throughout): fields stored in FP32 as deviations from the
 * liquid at rest (h_q - w_q, C - 1; g, P, U as they are), collision and U, P
 * in FP32, the fills, stream, prolong and restrict unchanged in form. Every
 * entry point works on it; host arrays stay FP64 in the reference layouts
 * (converted on the way in and out), and amrlbm_output reduces in FP64. */
int amrlbm_create_fp32(const amrlbm_params *p, amrlbm_mesh **out);
int amrlbm_destroy(amrlbm_mesh *m);
const char *amrlbm_last_error(void);

/* AMR_LBM_Rising_Bubble set-up (AMR_Rising_2D_OMP.f90:51-67): roots,
 * Initial_Mesh (:493-513), Initialize_AMR_Rising_Bubble (:142-231) and the
 * init-regrid loop, with bubble radius R centred at (x0,y0). */
int amrlbm_setup_rising_bubble(amrlbm_mesh *m, double R, double x0, double y0);
/* ninst independent copies of that set-up in one mesh: the nroots roots split
 * into ninst groups stacked in y, each group a rising-bubble domain of its
 * own with bounce-back walls at its ends and its bubble at (x0, y0) in its
 * own frame. Every step launch covers the leaves of all copies. No
 * counterpart in the reference (one domain per program); it is the
 * replicated-instance workload for measurements at a large working set.
 * ninst = 1 is amrlbm_setup_rising_bubble. */
int amrlbm_setup_replicas(amrlbm_mesh *m, int ninst, double R, double x0, double y0);
/* Device-resident set-up of n bubbles in one domain (radius R[k], centre
 * (x0[k], y0[k]), finest cells): the same roots, Initial_Mesh and
 * init-regrid loop as amrlbm_setup_rising_bubble, with C, its ghost rings and
 * the initial fields evaluated on the GPU, C = 0.5 + 0.5 tanh(2 d / W), d the
 * distance to the nearest bubble surface. No counterpart in the reference
 * (one bubble, profile evaluated on the host: amrlbm_setup_rising_bubble keeps
 * that, with the host's tanh, for bit-exact parity). */
int amrlbm_setup_bubbles(amrlbm_mesh *m, int n, const double *R, const double *x0, const double *y0);

/* One rising-bubble instance of an ensemble: its domain of two roots stacked
 * in y (L0 x 2 L0 finest cells, L0 = nx 2^max_levels, so its finest cells
 * are 1), its level range, its fluid constants (as amrlbm_params) and its
 * bubble in its own frame. */
typedef struct {
    int L0, min_levels, max_levels;
    double M0, W, sigma, beta, kappa, taul, tauh, rhol, rhoh, gy;
    double R, x0, y0;
} amrlbm_instance;
/* n independent rising-bubble domains of uneven size, level range and fluid
 * in one mesh (nroots = 2 n; the mesh's max_levels at least every instance's,
 * its max_blocks their total): instance q has roots 2q+1 and 2q+2 with
 * bounce-back walls at its ends. Set-up as AMR_LBM_Rising_Bubble
 * (AMR_Rising_2D_OMP.f90:51-67) per instance: roots of cell size L0/nx,
 * Initial_Mesh to its min_levels, its phase field profile with its W (host,
 * as amrlbm_setup_rising_bubble), and its own max_levels - min_levels passes
 * of the init-regrid loop. Every launch of the step then covers the leaves
 * of all instances; each kernel takes a block's fluid constants from its
 * instance and each regrid its level range. Every instance evolves as it
 * would alone. No counterpart in the reference (one domain per program). */
int amrlbm_setup_ensemble(amrlbm_mesh *m, int n, const amrlbm_instance *inst);
/* Output (AMR_Rising_2D_OMP.f90:456-489) of each instance of an ensemble
 * mesh in its own frame, out[n][9] as amrlbm_output's. */
int amrlbm_output_instances(amrlbm_mesh *m, int n, double *out);

/* ---- state exchange with the caller's Blk(:) mirror ----------------------- */
/* counts: Nblks, nLeaf, LevelMin, LevelMax, KK, free-list length, step T */
int amrlbm_get_counts(amrlbm_mesh *m, int counts[7]);
/* the step counter T (module variable, Parameters.f90:6; Main_Loop's DO T,
 * AMR_Rising_2D_OMP.f90:302): amrlbm_step advances it, the phase entry points
 * leave it to the caller, as the reference's subroutines leave it to the loop.
 * The regrid cadence of amrlbm_step and the snapshot header read it. */
int amrlbm_set_step(amrlbm_mesh *m, int T);
/* topology of blocks 1..Nblks, 21 ints each: Level, dx, I, J, Parent, IsLeaf,
 * hasLeaf, Refin, Deref, child(1:4), nghbr(1:8)%n (Parameters.f90:26-35);
 * freeq receives the free-list ids head first (Parameters.f90:41-45) */
int amrlbm_get_topology(amrlbm_mesh *m, int *blocks, int *freeq);
/* replace the mesh by a caller-built one (same 21-int rows); neighbour values
 * are recomputed from the quadtree and checked against the given ones */
int amrlbm_set_topology(amrlbm_mesh *m, int nblks, const int *blocks, int kk, int nfree,
                        const int *freeq, int T);

/* Restart files in the layout of the reference dump harness
 * (oracle/ref_harness.f90): header, free list, every block's topology row
 * and, for leaves, h, g, C, P, Ux, Uy in the reference's array layouts (FP64;
 * an FP32 mesh converts). No counterpart in the reference's own program, whose
 * only output is the Tecplot writer; the format lets a run restart and be
 * exchanged with the reference harness's snapshots. An FP32 mesh writes its
 * stored deviations (h_q - w_q, C - 1) exactly, marked by header word 1 = 2;
 * load converts between the forms as needed. load replaces the mesh; the
 * file's nx, ny, L0 and level range must equal the mesh's. */
int amrlbm_save_snapshot(amrlbm_mesh *m, const char *path);
int amrlbm_load_snapshot(amrlbm_mesh *m, const char *path);
/* TecPlot_CellCenter (TecPlot_ASCII.f90:1-86): the reference's ASCII
 * FE-quadrilateral file of the leaves (Array_of_Leaf_Blocks order) at step T,
 * written to path instead of Rising_Bubble_2D_NN.dat: node X, Y / L0, cell
 * Ux, Uy, node C - 1/2 (mean of the four cells around the node, first ghost
 * ring included), cell P D / (2 sigma), with the reference's edit descriptors.
 * D is the bubble diameter (D2Q9_Mod's D = 2 R). The file is the reference's
 * byte for byte, except that the time T sqrt(gy / D) in the third line can
 * differ in its last digit where the Fortran runtime picks another of the
 * equally short decimal forms of the same double. */
int amrlbm_tecplot_cell_center(amrlbm_mesh *m, const char *path, double D);
/* arrays of one live block (NULL pointers skip a field) */
int amrlbm_put_block(amrlbm_mesh *m, int id, const double *h, const double *g, const double *C,
                     const double *P, const double *Ux, const double *Uy);
int amrlbm_get_block(amrlbm_mesh *m, int id, double *h, double *g, double *C, double *P,
                     double *Ux, double *Uy);

/* ---- whole-mesh phases of one step (Main_Loop, AMR_Rising_2D_OMP.f90:302-364) */
int amrlbm_determine_min_max_levels(amrlbm_mesh *m, int *levelmin, int *levelmax); /* :434-452 */
int amrlbm_fill_ghost_cells_c2(amrlbm_mesh *m);       /* Fill_Ghost_Cells_C2.f90:2-40 */
int amrlbm_collide(amrlbm_mesh *m);                   /* :391-399 -> LBM_Subroutines.f90:116-189 */
int amrlbm_fill_ghost_cells_hg(amrlbm_mesh *m);       /* Fill_Ghost_Cells_hg.f90:2-49 */
int amrlbm_stream(amrlbm_mesh *m);                    /* :410-425 -> LBM_Subroutines.f90:41-90 */
/* Phase_Field_LBM_h_g_AMR (:381-430): collide, hg fill, stream; phase times in ms */
int amrlbm_phase_field_lbm_h_g(amrlbm_mesh *m, double ms[3]);
int amrlbm_macroscopic_values_h(amrlbm_mesh *m);      /* :328-335 -> LBM_Subroutines.f90:193-201 */
int amrlbm_fill_ghost_cells_c(amrlbm_mesh *m);        /* Fill_Ghost_Cells_C.f90:2-40 */
int amrlbm_macroscopic_values_g(amrlbm_mesh *m);      /* :346-354 -> LBM_Subroutines.f90:205-240 */
/* Refine_Derefine_Mesh (Refine_Derefine.f90:2-14). As in the reference,
 * Err_Interface reads C's ghost ring, which the step's C fill writes and a
 * derefined block holds as zeros (Restrict.f90:23-30): a second regrid with
 * no step between refines every block just derefined again. Callers step
 * between regrids, as Main_Loop does. */
int amrlbm_refine_derefine_mesh(amrlbm_mesh *m, double refin_tol, double deref_tol);
/* The device's regrid decision at the current state without regridding
 * (no counterpart in the reference): out[0] = 1 iff Refine_Derefine_Mesh
 * would change the mesh, out[1] = 1 iff it would refine a block (the plain
 * criterion finds the interface on an under-refined block). */
int amrlbm_regrid_check(amrlbm_mesh *m, int out[2]);
/* Refine_Derefine_Mesh with the blocks ids[0..n-1] flagged for refinement in
 * addition to the plain criterion (the learned regrid's predictions; the 2:1
 * closure and the derefinement rules apply as usual). */
int amrlbm_refine_derefine_predicted(amrlbm_mesh *m, const int *ids, int n); /* Refine_Derefine.f90:2-14 */

/* nsteps whole steps (fused kernels, one HIP graph per mesh topology); the
 * regrid runs after step T when (T+1) % regrid_every == 0, with the
 * reference's tolerances 0.002 / 0.001 (AMR_Rising_2D_OMP.f90:276-277) */
int amrlbm_step(amrlbm_mesh *m, int nsteps, int regrid_every);

/* the same steps as amrlbm_step, launched one kernel at a time with a HIP
 * event after each launch on the mesh's stream; stats[3k..3k+2] = total device
 * ms, launches and leaf cells processed of kernel class k (0: C2 fill +
 * collision, 1: hg fill + stream + C = sum h, 2: C fill + U,P); stats[9..10] =
 * regrid ms and count. No counterpart in the reference: this is the
 * measurement hook for the roofline report. */
int amrlbm_profile_steps(amrlbm_mesh *m, int nsteps, int regrid_every, double stats[12]);

/* Average device time (ms) of one launch of each step kernel over every leaf,
 * from reps back-to-back launches between two HIP events on the mesh's
 * stream: out[0] C2 fill + U,P + collision, out[1] hg fill + stream + C = sum h,
 * out[2] C fill + U,P. The mesh state is saved before and restored after.
 * No counterpart in the reference: the measurement hook for the roofline. */
int amrlbm_time_kernels(amrlbm_mesh *m, int reps, double out[3]);
/* Launch variants that leave every result bit for bit the same (no
 * counterpart in the reference): knobs[0] = the leaves' launch order (0: a
 * Morton curve cut into 8 runs, run c on XCD c; 1: the Morton curve as it
 * runs; 2: ascending id; 3: as 0, each run's blocks with the most ghost
 * operand loads first), knobs[1] = hg fill operands in flight per thread
 * (0: 2 and 1, 1: 4 and 2, 2: 8 and 4 recipes in the first and later stages),
 * knobs[2] = how the steps between two regrids run (0: a HIP graph of two
 * kernels per step over every leaf; 1: one persistent segment kernel per
 * regrid interval, one workgroup per leaf resident for all its steps, two
 * grid barriers per step, when every leaf's workgroup fits on the GPU at
 * once, else as 0). A new mesh starts at {0, 1, 0}. */
int amrlbm_set_tuning(amrlbm_mesh *m, const int knobs[3]);
int amrlbm_get_tuning(amrlbm_mesh *m, int knobs[3]);
/* The autotuner: for each of the 24 variants, times reps runs of the steps
 * amrlbm_step runs between two regrids (regrid_every steps, or twice that
 * when odd; no regrid) from the mesh's current state, which it then
 * restores, and keeps the fastest; table[4 v .. 4 v + 3] = order, depth,
 * execution, microseconds per step (-1: the segment kernel does not fit). */
int amrlbm_autotune(amrlbm_mesh *m, int reps, int regrid_every, double table[96]);

/* Phase-field mass correction (no counterpart in the reference, whose path
 * is not mass-conserving: the coarse/fine ghost interpolation of its steps
 * and its Prolong move the mass of Output, AMR_Rising_2D_OMP.f90:478).
 * on = 1: from the current state, each instance's phase-field mass
 * S = sum over leaf cells of C dx^2 (Output's mass is rhol * area +
 * (rhoh - rhol) S) is kept in FP64 on the device at its budget: S now, plus
 * the change each later mesh change's restrict and prolong make (summed in
 * FP64 before and after them). At every segment end (every regrid interval,
 * and the end of every stepping call) S is folded in FP64 and put back to
 * the budget by dC = lambda max(0, 4 C (1 - C)) on the interface cells
 * (lambda from the fold): C += dC, h_q += Gamma_q(U) dC and g_q +=
 * (Gamma_q(U) - w_q) (rhoh - rhol) dC / 3, which keep the cell's velocity and
 * pressure, before U, P and Err_Interface of that state are computed. It removes the steps' mass change (FP32
 * rounding and the interpolation's own) and keeps the regrids'. Off (0) by
 * default and after any set-up, restart or amrlbm_set_topology; with it on
 * the persistent segment kernel is not used. Works on FP64 and FP32 meshes
 * (an FP64 mesh then no longer follows the reference bit for bit). */
int amrlbm_set_mass_correction(amrlbm_mesh *m, int on);
/* The budget of each instance (n = 1, or the ensemble's instance count):
 * out[4 q .. 4 q + 3] = target, S at the last correction before it was
 * applied, the interface weight W, lambda. AMRLBM_ERR_STATE when off. */
int amrlbm_get_mass_budget(amrlbm_mesh *m, int n, double *out);

/* FP64 diagnostics over the leaf interiors (Output, AMR_Rising_2D_OMP.f90:456-489):
 * out[0..8] = rho_min, rho_max, max|Uy|, mass (sum rho dx^2), bubble area
 * (sum (1-C) dx^2), bubble centroid x, y, bubble rise velocity, non-finite count.
 * Returns AMRLBM_ERR_NONFINITE when the count is non-zero (outputs still set). */
int amrlbm_output(amrlbm_mesh *m, double out[9]);
/* amrlbm_output without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The per-leaf
 * partials live in a file-scope device array of the library (one per GPU,
 * held until the process ends) of at most 32768 leaves; a larger mesh gets
 * AMRLBM_ERR_CAPACITY. Same results as amrlbm_output, bit for bit. Calls on
 * meshes of one GPU share that array: calls that overlap on one GPU race.
 * The mesh's device tables must be current (any step, phase or regrid leaves
 * them so), else AMRLBM_ERR_STATE. */
int amrlbm_output_noalloc(amrlbm_mesh *m, double out[9]);

/* block until all queued device work of the mesh is done */
int amrlbm_synchronize(amrlbm_mesh *m);
/* device time (ms) of the last amrlbm_step call, measured with HIP events */
int amrlbm_last_step_ms(amrlbm_mesh *m, double *ms);
/* regrid bookkeeping since amrlbm_create, set-up included (no reference counterpart;
 * instrumentation): out[0] device regrid decisions taken by amrlbm_step, out[1] of them that
 * changed the mesh, out[2] segments launched ahead of such a change and taken back, host wall
 * ms spent in out[3] Refine_Derefine_Mesh (excluding table upload), out[4] table upload
 * (including out[5] ghost-program build and upload), out[6] graph capture and instantiation or
 * update; out[7] graphs captured (built, or updated after a topology change); out[8] changes
 * whose Refine and Deref flags came with the device decision (the others were flagged on the
 * host from a fresh Err_Interface) */
int amrlbm_regrid_stats(amrlbm_mesh *m, double out[9]);

/* ---- learned regrid predictor (no counterpart in the reference) ----------
 * A 3-layer 3x3 convolutional network (16 channels, BF16 weights on MFMA,
 * FP32 accumulate) over one leaf block's phase field as a signed distance
 * to the interface, (W/4) ln(C/(1-C)) in cells clamped to +-20 and divided by
 * 16 (18 x 18, first ghost ring), and Ux, Uy (interior, times 10, zero
 * ring): input rows of 3 x 18 x 18 floats;
 * output two logits per block: its region refined / derefined k steps ahead.
 * weights: AMRLBM_CNN_NPARAM floats (conv1 16x27, b1, conv2 16x144, b2,
 * conv3 2x144, b3; row k of a 3x3 filter is tap * cin_count + cin, tap =
 * 3 dy + dx). These entry points share one stream and device buffers of the
 * process: call them from one thread. */
#define AMRLBM_CNN_NPARAM 3058
int amrlbm_cnn_forward(const float *x, int n, const float *weights, float *logits);
/* gradient (AMRLBM_CNN_NPARAM) and value of the weighted logistic loss
 * sum over samples and channels of -[pw y log p + (1-y) log(1-p)] / n */
int amrlbm_cnn_gradient(const float *x, const float *y, int n, const float *weights, float pos_weight, float *grad,
                        float *loss);
/* Adam (lr, 0.9, 0.999) over epochs of shuffled minibatches, on the GPU;
 * weights in/out; epoch_loss[e] = mean minibatch loss of epoch e (may be null) */
int amrlbm_cnn_train(const float *x, const float *y, int n, float *weights, int epochs, int batch, float lr,
                     float pos_weight, unsigned seed, float *epoch_loss);
/* the network input of every leaf of the mesh: x (nLeaf rows), ids (block ids) */
int amrlbm_cnn_features(amrlbm_mesh *m, float *x, int *ids);
/* The mesh's learned regrid predictor for amrlbm_step_learned: weights as
 * above (copied to the device), a leaf flagged when logit 0 > threshold and
 * its Err_Interface exceeds gate times the refinement tolerance (gate >= 0;
 * the interface's profile already reaches the block); weights NULL removes it. */
int amrlbm_set_regrid_network(amrlbm_mesh *m, const float *weights, double threshold, double gate);
/* nsteps steps under the learned regrid's cadence, all decisions on the
 * device: at every multiple of full_every steps (T counted as amrlbm_step
 * counts it) the plain criterion (Refine_Derefine.f90:18-54, :97-194) decides
 * whether Refine_Derefine_Mesh changes the mesh; between them it judges every
 * step's state, and a state with the interface on an under-refined block is
 * regridded at once (a miss). The network (amrlbm_set_regrid_network) never
 * changes the mesh on its own: every regrid that does run also refines the
 * leaves it flags, so that the blocks the interface is about to reach are
 * refined ahead of it. Segments run back to back as in amrlbm_step; the host
 * works only when the mesh changes. stats (accumulated): [0] full regrids
 * reached, [1] leaves refined on the network's flag alone, [2] misses,
 * [3] steps judged. */
int amrlbm_step_learned(amrlbm_mesh *m, int nsteps, int full_every, long long stats[4]);

/* ---- per-block kernels on caller (host) arrays, reference layouts --------- */
int amrlbm_collision_block(const amrlbm_params *p, int dx, const double *C, const double *P,
                           const double *Ux, const double *Uy, double *h, double *g);    /* LBM_Subroutines.f90:116-189 */
int amrlbm_macro_g_block(const amrlbm_params *p, int dx, const double *g, const double *C,
                         double *P, double *Ux, double *Uy);                             /* LBM_Subroutines.f90:205-240 */
int amrlbm_stream_block(const amrlbm_params *p, int dx, double *f);                      /* LBM_Subroutines.f90:41-90 */

#ifdef __cplusplus
}
#endif
#endif
