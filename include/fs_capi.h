/* C-ABI of the MI355X residual-and-implicit-update library (libflowsolver.so).
 *
 * Plain pointers and sizes only; every call returns an fs_status (0 = OK).
 * Device state is resident in HBM between calls; host arrays move only
 * through the explicit set/get calls.  Nothing here exits the process: the
 * reference's fatal conditions (printf + exit) become status codes and
 * fs_last_error() text.
 *
 * Each entry point names the reference interface it replaces
 * (shedsaw/flow-solver, file:line).  The reference passes GRID and SYS_MEM pointers
 * plus PARAMS by value; here the mesh and solver handles own the same data,
 * laid out 0-based (reference index minus one).
 */
#ifndef FS_CAPI_H
#define FS_CAPI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FS_OK = 0,
  FS_ERR_ARG = -1,     /* invalid argument or size */
  FS_ERR_IO = -2,      /* file could not be read/written */
  FS_ERR_MESH = -3,    /* mesh rejected by preprocessing */
  FS_ERR_HIP = -4,     /* HIP runtime error or no device */
  FS_ERR_STATE = -5,   /* non-physical state: the reference would exit() */
  FS_ERR_NOMEM = -6,   /* device memory exhausted or over the admission budget */
  FS_ERR_UNSUPPORTED = -7
} fs_status;

/* Run controls: the 24 fields of the reference PARAMS
 * (include/params.h:19-76) in the order params.C:44-217 reads them. */
typedef struct {
  int nsteps, iter, jufreq;
  double CFL;
  int ramp;
  double gamma, alpha, mach;
  int subits, unsteady, tacc;
  double mintime;
  int method, order, foits, soits, toits;
  int isrestart, limit, reorder, ic, p_or_cv, start_post_pro, post_pro_freq;
} fs_params;

typedef struct fs_mesh fs_mesh;
typedef struct fs_solver fs_solver;

const char* fs_last_error(void);
int fs_device_count(int* count);
/* Achievable HBM bandwidth of the current device (no counterpart in the
 * reference): a streaming read and a copy of `bytes` (at least 1 MiB; use
 * more than the 256 MB last-level cache), `reps` timed passes each, in GB/s
 * (the copy counts read and write bytes).  Allocates and frees its buffers. */
int fs_hbm_probe(int64_t bytes, int reps, double* read_gbs, double* copy_gbs);

/* ---- parameters (params.C:27-219 read_parameters) ---- */
void fs_params_default(fs_params* p);
int fs_params_read(const char* path, fs_params* p);
int fs_params_write(const char* path, const fs_params* p);

/* ---- meshes: reference grid format (mesh_io.C:31-304 read_mesh,
 *      mesh_io.C:318-395 write_mesh) and seeded generators ---- */
int fs_mesh_read(const char* grid_path, const char* bc_path, fs_mesh** out);
int fs_mesh_write(const fs_mesh* m, const char* grid_path, const char* bc_path);
int fs_mesh_create_bump(int nx, int ny, double jitter, uint64_t seed, fs_mesh** out);
/* A mesh from arrays (0-based): the contents of the grid file that
 * mesh_io.C:31-214 reads into GRID (x, c2n[Tri], c2n[Quad], bs, bbc). */
int fs_mesh_create(int nn, const double* x, int ntri, const int* tri, int nquad, const int* quad, int nb,
                   const int* bseg_off, const int* bseg, const int* bbc, fs_mesh** out);
int fs_mesh_create_airfoil(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed, fs_mesh** out);
/* The same O-mesh (nwall even) with its wall cut at the leading and trailing
 * edges into boundaries upper (upper_bc), lower (lower_bc), farfield (0):
 * codes 1/1 give polyline walls, 17/19 the analytic NACA 0012 surfaces
 * (curved_boundaries.C:28-56). */
int fs_mesh_create_airfoil_split(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed,
                                 int upper_bc, int lower_bc, fs_mesh** out);
/* Annulus r_in <= r <= r_out: ntheta (even) nodes per ring on nrad
 * geometrically spaced rings, nodes at angles 0 and pi exactly on the x
 * axis (the reference's circle code assumes no segment crosses it,
 * curved_boundaries.C:152-154); boundaries inner (bc_in) and outer (bc_out).  Ring codes: 0-10, 11/12 at
 * radius 1, 21/22 at radius 300.  Codes 13-16 cannot close a ring: the
 * reference's arclength for them has no branch-cut handling
 * (curved_boundaries.C:1063-1093), so the segment next to the angle pi would
 * be about 2 pi R long. */
int fs_mesh_create_annulus(int ntheta, int nrad, double r_in, double r_out, double jitter, uint64_t seed, int bc_in,
                           int bc_out, fs_mesh** out);

/* ---- analytic curved walls, boundary codes 11-22 (curved_boundaries.C:28-44;
 *      include/curved_boundaries.h:21-33).  Points are (x, y).  The NACA
 *      arclength is the reference's integral evaluated by adaptive
 *      Gauss-Legendre quadrature after x = t^2 (the reference uses GSL QAGS
 *      at 1e-8 tolerance): NACA results agree with the reference's to that
 *      tolerance, circle results follow its formulas operation by
 *      operation. ---- */
/* curved_boundaries.C:91 curved_boundary_midpoint */
int fs_curved_midpoint(int bc, const double* xL, const double* xR, double* xM);
/* curved_boundaries.C:345 curved_boundary_gauss_points: GP[6] = the arc's
 * Gauss points (middle, +sqrt(3/5), -sqrt(3/5)) */
int fs_curved_gauss_points(int bc, const double* xL, const double* xR, double* GP);
/* curved_boundaries.C:701 curved_boundary_normal_vector */
int fs_curved_normal(int bc, const double* x, double* nx, double* ny);
/* curved_boundaries.C:1009 curved_boundary_arclength */
int fs_curved_arclength(int bc, const double* xL, const double* xR, double* len);
/* Preprocessing of main.C:149-559: Cuthill_McKee (reorder.C:58), metrics2D
 * (metrics.C:31), CSN/NSN/ESN/c2e maps (maps.C:35-517), Build_Edge_Data and
 * Build_Boundary_Edge_Data (geometry.C:119,407), Generate_Gaussian_Data_For_Edge
 * (geometry.C:993), cv_calc_area_Green / cv_calc_Moments (cv_calc.C:34,981),
 * Compute_Edge_Weights (gradient.C:923), Build_CRS (init_memory.C:36), the
 * order-3 stencils (solve.C:57-114 with maps.C:1316 Add_Layer_to_GNSN) and
 * the k=2 pseudo-inverse (compute_derivatives.C:859-940); plus the
 * multicolor row coloring used by the device Gauss-Seidel. */
int fs_mesh_prepare(fs_mesh* m, const fs_params* p);
/* Restart files (mesh_io.C:318-395 write_mesh, mesh_io.C:217-263 read_mesh
 * with isrestart = 1): the prepared mesh in its solver numbering plus the
 * last step, last CFL and the state of the nn real nodes; with_ghosts appends
 * the ng ghost states (an extension the reference's reader skips), which
 * makes a continuation bit-identical to an uninterrupted run.  Q holds
 * (nn+ng)*4 values as fs_solver_get(FS_F_Q).  Read a restart file with
 * fs_mesh_read and prepare it with isrestart = 1 (no renumbering). */
int fs_restart_write(const fs_mesh* m, const char* grid_path, const double* Q, int64_t n, int citer, double cfl,
                     int with_ghosts);
int fs_mesh_restart_info(const fs_mesh* m, int* has_restart, int* citer, double* cfl, int* has_ghosts);

typedef struct {
  int nn, ng, ne, nsub, nbe, nnz, ntri, nquad, nb, ncolor, nstencil;
  int prepared, order_prepared;
} fs_mesh_info;
int fs_mesh_get_info(const fs_mesh* m, fs_mesh_info* info);

typedef enum {
  FS_M_X = 0,          /* f64 2*(nn+ng) node and ghost coordinates */
  FS_M_TRI,            /* i32 3*ntri */
  FS_M_NN_MAP,         /* i32 nn file order -> new id (empty without reorder) */
  FS_M_NCSN, FS_M_CSN, /* i32 */
  FS_M_NNSN, FS_M_NSN, FS_M_ESN,
  FS_M_EDGES,          /* i32 2*ne */
  FS_M_C2E,            /* i32 */
  FS_M_SUB_EDGE, FS_M_SUB_ELEM,       /* i32 nsub */
  FS_M_SUB_N,          /* f64 3*nsub */
  FS_M_SUB_XM,         /* f64 2*nsub */
  FS_M_SUB_XG,         /* f64 6*nsub */
  FS_M_EL_CENT,        /* f64 2*nel */
  FS_M_BEDGE,          /* i32 5*nbe */
  FS_M_BEDGE_N,        /* f64 3*nbe */
  FS_M_BEDGE_XG,       /* f64 6*nbe: Gauss points of each bedge's half segment or half arc */
  FS_M_CV_AREA,        /* f64 nn */
  FS_M_MOMENTS,        /* f64 9*nn */
  FS_M_EDGE_WEIGHT,    /* f64 4*ne */
  FS_M_IA, FS_M_IAU, FS_M_JA,         /* i32 */
  FS_M_COLOR, FS_M_COLOR_OFF, FS_M_COLOR_ROWS, /* i32 */
  FS_M_ST_OFF, FS_M_ST_NODES,         /* i32 */
  FS_M_ST_W,           /* f64 5 per stencil entry */
  FS_M_BBC,            /* i32 nb */
  FS_M_BSEG_OFF, FS_M_BSEG, /* i32 */
  FS_M_BEDGE_GN,       /* f64 6*nbe: unit normal at each bedge Gauss point
                          (the curve's on a curved wall, the bedge normal otherwise) */
  FS_M_QUAD,           /* i32 4*nquad */
  FS_M_COUNT
} fs_mesh_array;
/* Borrowed host pointer to a preprocessed array; dtype 0 = int32, 1 = f64. */
int fs_mesh_get_array(const fs_mesh* m, int which, const void** data, int64_t* count, int* dtype);
void fs_mesh_destroy(fs_mesh* m);

/* ---- solver on one GPU ----
 * precision FS_PREC_F64: everything in FP64, the reference's algorithms
 * (finite-difference Jacobian, Gauss-Seidel sweeps) on multicolor order.
 * precision FS_PREC_MIXED: the analytic Jacobian (jacobian.C:301-658) and
 * the Gauss-Seidel sweeps in FP32 inside an FP64 defect-correction loop: the
 * residual (flux, reconstruction, limiter), the norms and the update stay
 * FP64, so the converged state is the FP64 residual's.
 * precision FS_PREC_MIXED16: FS_PREC_MIXED with the sweeps' off-diagonal
 * operands (the Jacobian blocks pre-multiplied by their row's FP64 diagonal
 * inverse) stored in FP16; the sweeps still compute in FP32 and everything
 * else is as in FS_PREC_MIXED, so only the linear solve's iterates move (by
 * the operands' rounding) and the converged state stays the FP64
 * residual's.  An operand outside the FP16 range fails the instance. */
enum { FS_PREC_F64 = 0, FS_PREC_MIXED = 1, FS_PREC_MIXED16 = 2 };
int fs_solver_create(const fs_mesh* m, const fs_params* p, int precision, fs_solver** out);
void fs_solver_destroy(fs_solver* s);
/* init_pv with ic_0 freestream, ghosts copied (ic.C:36-99, ic.C:808-913) */
int fs_solver_init_freestream(fs_solver* s);
/* Q and nQ from the solver mesh's restart payload (main.C:604-620: ghosts
 * copy their real node unless the file carries the ghost states). */
int fs_solver_init_restart(fs_solver* s);
typedef enum {
  FS_F_Q = 0,   /* f64 (nn+ng)*4: state at the start of the step, ghosts included */
  FS_F_NQ,      /* f64 (nn+ng)*4: Newton iterate */
  FS_F_DQ,      /* f64 nn*4 */
  FS_F_R,       /* f64 nn*4 */
  FS_F_RHS,     /* f64 nn*4 */
  FS_F_GRAD,    /* f64 nn*8  [n][var][x,y] */
  FS_F_PHI,     /* f64 nn     limiter scalar per node */
  FS_F_HESS,    /* f64 nn*20 [n][var][x,y,xx,yy,xy] */
  FS_F_PV,      /* f64 nn*4 point values */
  FS_F_EVCON,   /* f64 nn */
  FS_F_DT,      /* f64 nn */
  FS_F_LHS,     /* f64 nnz*16 row-major 4x4 blocks */
  FS_F_LHS32,   /* mixed precision only: the FP32 blocks, widened to f64, nnz*16 (with the
                   pre-multiplied sweep form, PARAMS.subits >= 30, the off-diagonal blocks
                   after a step are those products, as the sweeps read them) */
  FS_F_COUNT
} fs_field;
/* host <-> device copies (synchronous); n is the element count */
int fs_solver_set(fs_solver* s, int field, const double* host, int64_t n);
int fs_solver_get(fs_solver* s, int field, double* host, int64_t n);

/* Reference-function mirrors, each one device pass over the resident state.
 * ts is the pseudo-time step (grid->citer in the reference). */
int fs_build_matrix(fs_solver* s, int ts);                      /* init_memory.C:159 build_matrix */
int fs_gradient(fs_solver* s);                                  /* gradient.C:245 Compute_Gradient_LeastSquares */
/* cv_calc.C:790 cv_calc_eigenvalue_contribution: ev_con of the resident nQ
 * (sum over each node's faces of max|theta +- c| len, in the reference's
 * ascending face order); FS_ERR_STATE where the reference would exit */
int fs_eigenvalue_contribution(fs_solver* s);
/* cvbc.C:37 Set_Boundary_Values: the ghost states of the resident nQ from
 * their real nodes and their previous values (ghosts have memory);
 * FS_ERR_STATE where the reference would exit (cvbc.C:330-382) */
int fs_set_boundary_values(fs_solver* s);
int fs_derivatives(fs_solver* s);                               /* compute_derivatives.C:40 Compute_Derivatives(.,3) */
/* limiter.C:39 compute_limiter_barth, or limiter.C:511 compute_limiter_venkatakrishnan
 * when the solver's PARAMS has limit = 2 */
int fs_limiter(fs_solver* s, int ts);
int fs_interior_residual(fs_solver* s, int jupdate, int ts);    /* residual.C:40 build_interior_residuals */
int fs_boundary_residual(fs_solver* s, int jupdate);            /* residual.C:359 build_boundary_residuals */
int fs_linear_solve(fs_solver* s, double* rms);                 /* linear_algebra.C:122 linear_solve */
/* solve.C:44 solve(): one pseudo-time step; returns ||dQ||_2 of the last
 * Newton iteration and sum(R^2) of the last residual. */
int fs_solve(fs_solver* s, int ts, double* dq_norm, double* r2);
/* Batched Roe_flux_centered (flux.C:299): n faces, device pointers;
 * normals[3*i] = (nx, ny, len), qL/qR/flux 4 per face. */
int fs_roe_flux_batch(int n, double gamma, const double* d_normals, const double* d_qL, const double* d_qR,
                      double* d_flux, void* stream);
/* The same with a choice of arithmetic: form 0 is the reference's operation
 * order (bit-exact, as fs_roe_flux_batch); form 1 the characteristic-wave form
 * the mixed path's FP64 residual uses (same flux to a few ulps of |F|). */
int fs_roe_flux_batch_form(int n, int form, double gamma, const double* d_normals, const double* d_qL,
                           const double* d_qR, double* d_flux, void* stream);
/* solve.C:785-793 and solve.C:885-893, the norm loops of solve(): over nn
 * nodes of device arrays d_R and d_dQ (4 values per node),
 * d_out[0] = sum R^2 and d_out[1] = sum dQ^2 (d_out a device pointer to 2
 * doubles).  Summed in the solver's fixed order (chunks of 1024 nodes), so
 * the same arrays give the same bits as fs_solve's r2.  Allocates its chunk
 * partials and waits for `stream` before returning. */
int fs_state_norms(int64_t nn, const double* d_R, const double* d_dQ, double* d_out, void* stream);
/* No counterpart in the reference.  fs_state_norms for callers that cannot
 * allocate device memory (code under HIP stream capture, serving loops that
 * must not call hipMalloc): same arguments and the same result bits, but the
 * chunk partials live in a fixed per-device buffer of the library (1 MiB,
 * held until the process ends), so nn is at most 67,108,864 (FS_ERR_ARG
 * beyond); nothing is allocated and the call only enqueues work on `stream`.
 * Every call on one device shares that buffer: calls that overlap on one
 * device (on different streams or threads) race. */
int fs_state_norms_noalloc(int64_t nn, const double* d_R, const double* d_dQ, double* d_out, void* stream);

/* grid_io.C:1390 Compute_Body_Forces: forces[0] (lift) and forces[1] (drag)
 * integrated over the walls (odd bc <= 10) from the resident state after
 * step `citer` (grid->citer selects the boundary reconstruction, as
 * reconstruct.C:1708 Reconstruct_Gauss_Boundary does). */
int fs_body_forces(fs_solver* s, int citer, double forces[2]);
/* grid_io.C:841 write_surface_cp, nodal values: for boundary b (0-based) the
 * x coordinate and Cp of its unique segment nodes in order of first
 * appearance (count 0 for a boundary that is not a wall).  With x or cp NULL
 * only *count is set. */
int fs_surface_cp(fs_solver* s, int b, double* x, double* cp, int64_t cap, int64_t* count);
/* the same written as the reference writes it: dir/surface_<b+1>_pressure.dat
 * for every wall, "%f  %f" lines in `sort -g` order */
int fs_write_surface_cp(fs_solver* s, const char* dir);

/* Steady solve to ||R||/||R0|| <= tol (R0 = residual at the first step of the
 * highest-order phase) or max_steps; returns steps taken and the drop. */
int fs_solve_steady(fs_solver* s, int max_steps, double tol, int* steps, double* drop);

/* ---- timing helpers for the benchmark: device-side event timing of the
 *      pseudo-time step and of the dominant kernel ---- */
typedef struct {
  double step_ms;        /* mean wall time per pseudo-time step */
  double edge_ms;        /* mean device time of the face kernel: flux + FD Jacobian
                            (FP64); the boundary-edge kernel (mixed) */
  double node_ms;        /* mean device time of the node kernel: assembly (FP64);
                            FP64 residual + FP32 Jacobian blocks + assembly (mixed) */
  double gs_ms;          /* mean device time of the Gauss-Seidel sweep launches */
  double lu_ms;          /* mean device time of the diagonal-block factorisation */
  double other_ms;
  int64_t flux_evals;    /* residual Roe evaluations per step (subedges+bedges) */
} fs_timing;
int fs_solver_time_steps(fs_solver* s, int ts_first, int nsteps, fs_timing* t);

/* Run fs_solve's pseudo-time step as one captured hipGraph launch (on != 0)
 * or as individually enqueued kernels (0, the default).  No counterpart in
 * the reference (its step is a sequence of CPU loops). */
int fs_solver_use_graph(fs_solver* s, int on);

/* Launch-shape autotuning (no counterpart in the reference).  Times every
 * candidate workgroup size / lanes-per-row / sweep prefetch depth / kernel
 * form of the tunable kernels on the handle's own data (after at least one
 * step) and keeps the fastest; none of the choices changes any result.  The choice is
 * appended to the cache file (FS_TUNE_CACHE, default
 * $HOME/.cache/flowsolver/tune_cache.txt) under (arch, precision, k=2
 * present, log2 node count), and handles created later with the same key
 * start from it.  values: gs_block, gs_lanes, gs32_block, gs32_pre,
 * node_block, face_block, cell_block, deriv_lds (k=2 derivatives with the
 * neighbour states staged in LDS per 32-node tile: 1, gathered through L2:
 * 0), gs32_scaled (FP32 sweeps on blocks and right-hand sides pre-multiplied
 * by the diagonal inverse once per Newton iteration: 1, applying the stored
 * inverse in every sweep: 0), node_fused (pre-multiplied operands formed by
 * the mixed node pass itself, with each row's blocks staged in LDS: 1, or by
 * a separate pass over the stored blocks: 0; the same bits either way),
 * gs32_dpp (the 4-lane sweeps: each lane loads one component of a
 * neighbour's dQ and the row's 4 lanes exchange them: 1, or every lane loads
 * all four: 0).  gs32_pre and gs32_dpp apply to the FP64 4-lane sweep too;
 * gs32_pre 0 (FP64: no prefetch, blocks loaded one at a time) is a candidate
 * there. 
 * gs32_scaled is not tuned or cached: it is 1 when PARAMS.subits >= 30 (and
 * always with FS_PREC_MIXED16) and can be pinned with the setter; the two
 * forms round differently in FP32, so it moves the mixed path's iterates by
 * FP32 rounding, never its converged state. */
#define FS_TUNING_VALUES 11
int fs_solver_autotune(fs_solver* s, double* seconds);
int fs_solver_get_tuning(const fs_solver* s, int* values, int* from_cache);
/* sets the launch shapes (FS_TUNING_VALUES values in the order above; each
 * must be one of the autotuner's candidates); results do not change */
int fs_solver_set_tuning(fs_solver* s, const int* values);

/* ---- batches: many independent steady solves (solve.C:44 each) packed
 *      CSR-style into one device image; every kernel of a pseudo-time step
 *      is one launch over all instances, optionally one hipGraph launch per
 *      step.  Instances keep their own mesh, PARAMS and step counter
 *      (grid->citer); they share subits, iter and jufreq.  An instance's
 *      results do not depend on the other instances of the batch.  No
 *      counterpart in the reference, which solves one mesh per process. ---- */
typedef struct fs_batch fs_batch;
typedef enum {
  FS_INST_RUNNING = 0,
  FS_INST_CONVERGED = 1,   /* residual drop reached the tolerance */
  FS_INST_MAX_STEPS = 2,   /* step limit reached first */
  FS_INST_FAILED = 3,      /* non-physical state (after its recoveries, if enabled) */
  FS_INST_PAUSED = 4,      /* masked off by fs_batch_set_active */
  FS_INST_CANCELLED = 5    /* serving: cancelled before it finished */
} fs_inst_status;
typedef struct {
  int ninst, nn, ne, nsub, nbe, nnz, ncolor;
  int64_t device_bytes;
} fs_batch_info;
/* meshes prepared (and kept alive by the caller until fs_batch_destroy) */
int fs_batch_create(int n, const fs_mesh* const* meshes, const fs_params* params, int precision, fs_batch** out);
void fs_batch_destroy(fs_batch* b);
int fs_batch_get_info(const fs_batch* b, fs_batch_info* info);
int fs_batch_init_freestream(fs_batch* b);
int fs_batch_use_graph(fs_batch* b, int on);
/* Deferred k=2 fields (on by default; no counterpart in the reference).  In
 * the k=2 phase no residual or Jacobian reads the k=2 derivatives, point
 * values or limiter values (residual.C:208-290 overwrites the Gauss-point
 * flux with the nodal one; the boundary Gauss branch, residual.C:446, is
 * disabled), which solve.C:396-472 still evaluates every step.  With
 * deferral a step skips them and keeps the state it started from, and every
 * output that reads them (body forces, field reads, single passes) first
 * evaluates them from that state: all results are those of evaluation
 * every step, bit for bit (instances with one Newton iteration per step and
 * no limiter freeze; others evaluate them every step).  on = 0 evaluates
 * them every step. */
int fs_batch_defer_k2(fs_batch* b, int on);
int fs_batch_autotune(fs_batch* b, double* seconds);
int fs_batch_get_tuning(const fs_batch* b, int* values, int* from_cache);
int fs_batch_set_tuning(fs_batch* b, const int* values);
/* Re-packing during fs_batch_solve_steady (off by default; no counterpart in
 * the reference): once at most half of the newest engine's instances still
 * run (and it holds at least 8), the running ones move to a new, smaller
 * engine on the device, so later launches cover only live work.  Results do
 * not change; stopped instances keep their final fields.  Each re-pack
 * builds a new device image on the host, which costs more than it saves on
 * the 256-solve workload (measured), so it pays only for long tails of a
 * few running solves. */
int fs_batch_set_compaction(fs_batch* b, int on);
/* Recovery from non-physical states (off by default: max_recoveries 0; no
 * counterpart in the reference, which exits, cvbc.C:330-382 and
 * init_memory.C:257-268).  With max_recoveries > 0 every instance keeps two
 * checkpoints of its state with its ghosts in device memory (64 bytes per
 * node and ghost): the latest, taken after its first healthy step and then
 * every checkpoint_every steps unless its residual is above ten times the
 * best of its current phase, and the start of its current phase, taken
 * after the last step of each phase.  A step that sets a device error flag
 * or gives a non-finite residual returns the instance to its latest
 * checkpoint, step counter included, with half its CFL target and the CFL
 * ramp restarted from 1; after max_recoveries such restores the next
 * failure is final (FS_INST_FAILED).  With limiter_fallback > 0 a failure of
 * an instance that runs without a limiter (PARAMS.limit 0) in the phase
 * where a limiter would run instead restarts that phase from its start with
 * the Barth limiter switched on, frozen (fs_batch_set_limiter_freeze) after
 * limiter_fallback limited steps: such an instance converges to the steady
 * state of another discretization, one that depends on its iterate path.
 * With first_order_fallback != 0 the next failure of an instance in a
 * reconstructing phase (linear, or k=2) restarts that phase from its start
 * with first-order face states for the rest of its solve, at its own CFL
 * target again (the halvings of the earlier restores are undone): it
 * converges to the first-order steady state.  Instances without a failure follow
 * exactly the path they follow with recovery off.  Steps a restore rolled
 * back still count in `steps` and against max_steps.  Set it before the
 * first step. */
int fs_batch_set_recovery(fs_batch* b, int max_recoveries, int checkpoint_every, int limiter_fallback,
                          int first_order_fallback);
/* Stall test of the recovery policy (off by default: 0; no counterpart in
 * the reference).  A healthy instance whose residual norm has not fallen
 * below half its lowest mark of the current phase for
 * max(100, ceil(steps_per_sqrt_node * sqrt(nodes))) steps has stalled: the
 * stall starts the next fallback of fs_batch_set_recovery that applies to it
 * (limiter, then first order; it counts as a restore) at the CFL it runs
 * with.  With no fallback left the instance runs on.  Instances that do not
 * stall follow exactly the path they follow with the test off. */
int fs_batch_set_stall(fs_batch* b, double steps_per_sqrt_node);
/* the fallbacks that acted on a solve (bits; FS_FALLBACK_STALL: a stall
 * started at least one of them) */
enum { FS_FALLBACK_LIMITER = 1, FS_FALLBACK_FIRST_ORDER = 2, FS_FALLBACK_STALL = 4 };
/* restores instance i has used and the FS_FALLBACK_* bits of the fallbacks
 * that acted on it (either pointer may be NULL) */
int fs_batch_inst_recovery(const fs_batch* b, int i, int* restores, int* fallbacks);
/* Limiter freeze (off by default: steps 0; no counterpart in the reference,
 * whose limiter runs every step, solve.C:440-471): once an instance's limited
 * phase has run more than `steps` steps, the limiter pass keeps its values
 * (phi) from then on.  The limited iteration then converges to the steady
 * state of the discretization with those frozen limiter values instead of
 * cycling; which values are frozen depends on the iterate path (precision,
 * CFL schedule), and so does that steady state.  The stopping test is
 * unchanged. */
int fs_batch_set_limiter_freeze(fs_batch* b, int steps);
/* convergence mask: a masked-off instance is skipped by every kernel */
int fs_batch_set_active(fs_batch* b, int i, int active);
/* one pseudo-time step of every active instance (each step counter advances);
 * r2[i] = sum R^2, dq[i] = ||dQ||_2 of instance i (either may be NULL) */
int fs_batch_step(fs_batch* b, double* r2, double* dq);
/* steady solves: each instance stops on its own when ||R||/||R0|| <= tol (R0
 * as fs_solve_steady) and is masked off; runs until every instance stopped
 * or max_steps steps.  steps/drop/status: one entry per instance (NULL ok). */
int fs_batch_solve_steady(fs_batch* b, int max_steps, double tol, int* steps, double* drop, int* status);
int fs_batch_inst_status(const fs_batch* b, int i, int* status, int* steps, double* drop);
int fs_batch_get(fs_batch* b, int i, int field, double* host, int64_t n);
int fs_batch_set(fs_batch* b, int i, int field, const double* host, int64_t n);
/* grid_io.C:1390 Compute_Body_Forces of every instance from its state after
 * its last step: forces[2*i] lift, forces[2*i+1] drag */
int fs_batch_body_forces(fs_batch* b, double* forces);
/* nsteps timed pseudo-time steps of the active instances (step counters
 * advance).  Without graphs (fs_batch_use_graph off) the kernel fields of
 * fs_timing are filled from events on the engines' streams; with graphs only
 * step_ms and flux_evals. */
int fs_batch_time_steps(fs_batch* b, int nsteps, fs_timing* t);
/* device bytes a mesh costs inside a batch/solver of the given precision */
int fs_mesh_footprint(const fs_mesh* m, int precision, int64_t* bytes);

/* ---- serving: a request loop that accepts steady-solve requests (a mesh
 *      and its PARAMS: Mach, angle of attack, order ...), prepares meshes on
 *      host worker threads, admits prepared requests at pseudo-time-step
 *      boundaries while their device footprint fits the budget, keeps every
 *      admitted solve resident in HBM, and overlaps the steps of all resident
 *      solves on the device (requests admitted together share one batched
 *      engine; each engine runs on its own stream).  A solve ends when
 *      ||R||/||R0|| <= tol (as fs_solve_steady), at its step limit, or on a
 *      non-physical state that its recoveries (if enabled) did not cure.
 *      No counterpart in the reference. ---- */
typedef struct fs_server fs_server;
typedef struct {
  int64_t device_budget;   /* bytes of device memory for resident solves (0: 80% of free) */
  int precision;           /* FS_PREC_F64, FS_PREC_MIXED or FS_PREC_MIXED16 */
  int max_steps;           /* step limit of a request (see max_steps_per_sqrt_node) */
  double tol;
  int prep_threads;        /* host preprocessing workers */
  int use_graph;           /* one hipGraph launch per engine step */
  int max_recoveries;      /* recovery policy of fs_batch_set_recovery (0: off) */
  int checkpoint_every;
  double max_steps_per_sqrt_node;  /* > 0: a request's limit is max(max_steps,
                                      ceil(this * sqrt(real nodes))) */
  int limiter_freeze;      /* fs_batch_set_limiter_freeze's steps (0: off) */
  int limiter_fallback;    /* fs_batch_set_recovery's limiter_fallback (0: off) */
  int first_order_fallback;  /* fs_batch_set_recovery's first_order_fallback (0: off) */
  double stall_steps_per_sqrt_node;  /* fs_batch_set_stall's test (0: off) */
} fs_server_config;
typedef struct {
  int64_t id;
  int status;              /* fs_inst_status; FS_INST_FAILED also for requests rejected at admission */
  int steps;
  double drop;
  double t_submit, t_admit, t_done;   /* seconds on the server clock (fs_server_clock) */
  int64_t device_bytes;
  int nn, ng;
  double forces[2];        /* grid_io.C:1390 Compute_Body_Forces of the final state (lift, drag) */
  int recoveries;          /* restores the recovery policy made */
  int fallbacks;           /* FS_FALLBACK_* bits of the recovery policy's fallbacks that acted */
  int max_steps;           /* the request's step limit */
  double r0;               /* ||R|| at the reference step: drop = ||R|| / r0 (-1 before it) */
} fs_request_info;
void fs_server_config_default(fs_server_config* c);
int fs_server_create(const fs_server_config* c, fs_server** out);
/* stops the loop; outstanding requests are dropped */
void fs_server_destroy(fs_server* s);
/* takes ownership of `mesh` (prepared or not; it is prepared with `p` on a worker) */
int fs_server_submit(fs_server* s, fs_mesh* mesh, const fs_params* p, int64_t* id);
/* waits up to timeout_s for the next finished request; *got = 0 on timeout */
int fs_server_wait(fs_server* s, double timeout_s, fs_request_info* out, int* got);
/* final state of a finished request: (nn+ng)*4 values, ghosts after the real nodes */
int fs_server_get_state(fs_server* s, int64_t id, double* Q, int64_t n);
/* forgets a finished request (its state is freed) */
int fs_server_release(fs_server* s, int64_t id);
int fs_server_cancel(fs_server* s, int64_t id);
int fs_server_counts(fs_server* s, int* queued, int* running, int* finished);
double fs_server_clock(fs_server* s);

#ifdef __cplusplus
}
#endif
#endif
