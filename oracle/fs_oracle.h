/* Oracle: a plain-C, single-threaded restatement of the reference's FP64
 * residual-and-implicit-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline; the product never calls it.
 *
 * Arrays are 0-based (reference index minus one) and laid out like the
 * reference's GRID/SYS_MEM arrays (include/grid.h:19-307,
 * include/sys_mem.h:19-50).  Loop orders and arithmetic order follow the
 * reference expression by expression so results agree bit for bit with the
 * reference build where the same operations run.
 *
 * Parity status per function (how each is pinned):
 *   fso_roe_flux            golden vectors from the reference build; consistency,
 *                           conservation, supersonic upwinding (tests)
 *   fso_build_matrix        reference build step histories (ev_con, dt, LHS diag)
 *   fso_gradient            reference build; exact for linear fields
 *   fso_limiter_barth       reference build (phi per step)
 *   fso_limiter_venkat      reference build (phi per step, k=1 face points)
 *   fso_interior_residual   reference build (R, LHS per step)
 *   fso_boundary_residual   reference build (R, LHS, ghost states per step)
 *   fso_linear_solve        reference build (dQ per step); dense solve on tiny systems
 *   fso_solve               reference build step histories
 *   analytic Jacobian       (fso_set_analytic_jacobian) restates jacobian.C:301-658, which the
 *                           reference's solve never calls: bit for bit against the reference
 *                           build's own compute_approximate_jacobian_roe over every face at
 *                           unequal reconstructed face states (oracle/ref_pointwise jac mode);
 *                           equal to the forward difference where the face states are equal
 *   fso_build_pinv          numpy pinv and quadratic exactness (the reference build
 *                           cannot run the k=2 factorization: LAPACK is absent)
 *   fso_derivatives         quadratic exactness of hess/point_value (tests);
 *                           no reference-build vector exists for the k=2 phase
 *   fso_body_forces         reference build (forces per step, first-order and linear
 *                           phases); k=2 phase: quadratic exactness of the Gauss-point
 *                           reconstruction and its reduction to the linear phase (tests);
 *                           curved walls (codes 11-22, Gauss points and normals read from
 *                           the mesh): no net force from a uniform state on a closed
 *                           curved wall (the reference's curved code needs GSL and does
 *                           not build here)
 *   fso_surface_cp          closed form (uniform state gives Cp of the freestream pressure)
 *   fso_limiter_barth, fso_limiter_venkat (k=2 Gauss points): from CV averages of
 *                           quadratic fields, phi equals the per-point function at the
 *                           exact field values of every Gauss point (tests); Barth also
 *                           by linear preservation and a bowl limited only near its
 *                           minimum; no reference-build vector (it cannot enter the k=2
 *                           phase here)
 *   k=2 phase of fso_solve  no reference-build vector; pinned by the invariants of its
 *                           parts (first-order residual and Jacobian as in the reference's
 *                           k=2 phase, derivatives, limiters, body forces above)
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nn, ng, ne, nsub, nbe, nnz, nb, nel;
  const double* x;          /* 2*(nn+ng) */
  const int* edges;         /* 2*ne, L<R */
  const int* sub_edge;      /* nsub */
  const int* sub_elem;      /* nsub */
  const double* sub_n;      /* 3*nsub (nx, ny, len) */
  const double* sub_xm;     /* 2*nsub */
  const double* sub_xg;     /* 6*nsub */
  const double* el_cent;    /* 2*nel */
  const int* bedge;         /* 5*nbe (real, ghost, gelem, b, s) */
  const double* bedge_n;    /* 3*nbe */
  const int* bbc;           /* nb */
  const double* cv_area;    /* nn */
  const double* moments;    /* 9*nn */
  const double* edge_weight;/* 4*ne */
  const int* ia;            /* nn+1 */
  const int* iau;           /* nn */
  const int* ja;            /* nnz */
  const int* st_off;        /* nn+1 (k=2 only) */
  const int* st_nodes;      /* stencil without self (k=2 only) */
  int ncolor;
  const int* color_off;     /* ncolor+1 */
  const int* color_rows;    /* nn */
  const int* bseg_off;      /* nb+1: segments of boundary b */
  const int* bseg;          /* 2 per segment (n0, n1) */
  const double* bedge_xg;   /* 6*nbe: Gauss points of the bedge (read for curved walls only) */
  const double* bedge_gn;   /* 6*nbe: unit normal at each Gauss point (curved walls only) */
} fso_mesh;

typedef struct {
  int nsteps, iter, jufreq;
  double CFL;
  int ramp;
  double gamma, alpha, mach;
  int subits, unsteady, tacc;
  double mintime;
  int method, order, foits, soits, toits;
  int isrestart, limit, reorder, ic, p_or_cv, start_post_pro, post_pro_freq;
} fso_params;

typedef struct {
  double* Q;      /* (nn+ng)*4 */
  double* nQ;     /* (nn+ng)*4 */
  double* pQ;     /* (nn+ng)*4 */
  double* dQ;     /* nn*4 */
  double* R;      /* nn*4 */
  double* RHS;    /* nn*4 */
  double* grad;   /* nn*8   [n][var][dim] */
  double* phi;    /* nn*8   [n][var][phi,sigma] */
  double* hess;   /* nn*20  [n][var][5] */
  double* pv;     /* nn*4 point values */
  double* W;      /* 5*len(st_nodes) pseudo-inverse, weights folded (k=2) */
  double* ev_con; /* nn */
  double* dt;     /* nn */
  double* LHS;    /* nnz*16 */
  int citer;
  double cfl;
  double last_rms;   /* sum R^2 of the last Newton iteration */
  double last_lin;   /* RMS returned by the last linear solve */
} fso_state;

int fso_roe_flux(double nx, double ny, double len, double gamma, const double qleft[4], const double qright[4],
                 double flux[4]);
void fso_init_freestream(const fso_mesh* m, const fso_params* p, fso_state* s);
int fso_build_matrix(const fso_mesh* m, const fso_params* p, fso_state* s);
void fso_gradient(const fso_mesh* m, fso_state* s);
int fso_build_pinv(const fso_mesh* m, fso_state* s);
void fso_derivatives(const fso_mesh* m, fso_state* s);
void fso_limiter_barth(const fso_mesh* m, const fso_params* p, fso_state* s);
void fso_limiter_venkat(const fso_mesh* m, const fso_params* p, fso_state* s);
void fso_interior_residual(const fso_mesh* m, const fso_params* p, fso_state* s, int jupdate, int ts);
int fso_boundary_residual(const fso_mesh* m, const fso_params* p, fso_state* s, int jupdate);
/* multicolor = 0: the reference's sequential symmetric Gauss-Seidel;
 * multicolor = 1: the same sweeps with rows visited color by color. */
double fso_linear_solve(const fso_mesh* m, const fso_params* p, fso_state* s, int multicolor);
/* 1: the interior Jacobian is compute_approximate_jacobian_roe (jacobian.C:301-658)
 * instead of the forward difference; used to check the FP32 analytic blocks */
void fso_set_analytic_jacobian(int on);
double fso_solve(const fso_mesh* m, const fso_params* p, fso_state* s, int ts, int multicolor);
/* Compute_Body_Forces (grid_io.C:1390-1697), straight boundaries (bc <= 10):
 * forces[0] (lift) and forces[1] (drag) from the state after step `citer`. */
void fso_body_forces(const fso_mesh* m, const fso_params* p, const fso_state* s, int citer, double forces[2]);
/* write_surface_cp (grid_io.C:841-1375, the nodal branch): for boundary b
 * (0-based) the unique segment nodes in first-appearance order, their x and
 * Cp from Q.  Returns the count (0 for a boundary that is not a wall). */
int fso_surface_cp(const fso_mesh* m, const fso_params* p, const fso_state* s, int b, double* x, double* cp);

#ifdef __cplusplus
}
#endif
#endif
