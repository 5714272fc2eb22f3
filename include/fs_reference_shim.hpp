/* Reference-shaped entry points (C++, header only) over the C-ABI.
 *
 * The reference drives the residual-and-implicit-update path through free
 * functions that take GRID* / SYS_MEM* and PARAMS by value.  This header gives
 * those functions with the same names and signatures:
 *
 *   solve                              include/solve.h:21           (solve.C:44)
 *   build_matrix                       include/init_memory.h:24     (init_memory.C:159)
 *   build_interior_residuals           include/residual.h:21        (residual.C:40)
 *   build_boundary_residuals           include/residual.h:23        (residual.C:359)
 *   compute_limiter_barth              include/limiter.h:21         (limiter.C:39)
 *   compute_limiter_venkatakrishnan    include/limiter.h:23         (limiter.C:511)
 *   Compute_Gradient_LeastSquares      include/gradient.h:21        (gradient.C:245)
 *   Compute_Derivatives                include/compute_derivatives.h:19 (compute_derivatives.C:40)
 *   linear_solve                       include/linear_algebra.h:23  (linear_algebra.C:122)
 *   Roe_flux_centered                  include/flux.h:26            (flux.C:299)
 *   Reconstruct                        include/reconstruct.h:21     (reconstruct.C:38)
 *   Reconstruct_Gauss                  include/reconstruct.h:25     (reconstruct.C:1109)
 *   Reconstruct_Gauss_Boundary         include/reconstruct.h:27     (reconstruct.C:1708)
 *   compute_numerical_jacobian_roe     include/jacobian.h:21        (jacobian.C:45)
 *   compute_approximate_jacobian_roe   include/jacobian.h:28        (jacobian.C:301)
 *   compute_boundary_approximate_jacobian_roe include/jacobian.h:32 (jacobian.C:681)
 *   compute_boundary_numerical_jacobian_roe include/jacobian.h:25 (jacobian.C:205)
 *   Get_Boundary_Value                 include/cvbc.h:24            (cvbc.C:502)
 *   Set_Boundary_Values                include/cvbc.h:22            (cvbc.C:37)
 *   cv_calc_eigenvalue_contribution    include/cv_calc.h:28         (cv_calc.C:790)
 *   Compute_Edge_Weights               include/gradient.h:23        (gradient.C:923)
 *   Build_CRS                          include/init_memory.h:22     (init_memory.C:36)
 *
 * as templates over the caller's GRID / SYS_MEM / PARAMS types, reading the
 * fields by the reference's names and layouts (include/grid.h:224-306,
 * include/sys_mem.h:44-48, include/params.h:51-74): 1-based arrays
 * with slot 0 unused, Q/nQ interleaved (rho, rhou, rhov, E) with ghosts after
 * the nn real nodes, grad [node][var][x,y], hess [node][var][9 moments],
 * phi [node][var][phi,sigma], c2n[Tri]/c2n[Quad] connectivity, bs[b][s][2]
 * boundary segments.  The reference's own GRID, SYS_MEM and PARAMS classes
 * satisfy them, so a translation unit of the reference can define its
 * entry points by forwarding, e.g.
 *   double solve(GRID* g, SYS_MEM* s, PARAMS p, int ts) { return fsref::solve(g, s, p, ts); }
 * fsref::GRID / fsref::SYS_MEM below are plain structs with just those
 * fields, for callers without the reference's classes.
 *
 * The first call on a GRID builds its device-resident twin from x, c2n, bs
 * and bbc (the GRID is taken in its current numbering, i.e. after the
 * reference's own reordering, so preprocessing runs with reorder = 0).  Each
 * loop-level call copies the host fields the reference function reads to the
 * twin, runs the device pass and copies back the fields it writes; solve()
 * is the fast path (state in, one device step, state out).  Arrays the
 * reference allocates lazily (R, dQ, grad, phi, hess, point_value, ev_con,
 * dt; LHS and RHS of SYS_MEM) are calloc'ed here when null and belong to the
 * caller.  The reference's
 * fatal conditions (printf + exit) stay fatal: a failed call prints the
 * library's message and exits with status 1.
 *
 * Per-node values only: phi receives the node scalar in its [0] and [1]
 * entries (the ones the path consumes); hess entries 5..8 of each variable
 * are not produced (the k=2 path uses 0..4).
 *
 * The per-face functions (Roe_flux_centered, Reconstruct*, the four
 * Jacobian functions, Get_Boundary_Value) run on the host from the shared
 * per-face header the kernels are built from (include/fs_face.hpp) and read
 * or update the caller's own arrays, as the reference's do; Set_Boundary_Values and
 * cv_calc_eigenvalue_contribution are device passes over the twin;
 * Compute_Edge_Weights and Build_CRS copy the host preparation's arrays.
 * Reconstruction orders 0, 1 (k=1) and 3 (k=2) are supported; order 2 and 4
 * stop as an unsupported configuration.
 */
#ifndef FS_REFERENCE_SHIM_HPP
#define FS_REFERENCE_SHIM_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fs_capi.h"
#include "fs_face.hpp"

namespace fsref {

// element-type slots of num_elem / c2n (the reference names them Edge, Tri, Quad)
enum { kEdge = 0, kTri = 1, kQuad = 2, kNumElemType = 3 };

typedef fs_params PARAMS;  // the reference's 24 fields, same order (include/params.h:19-76)

struct Twin {
  fs_mesh* mesh = nullptr;
  fs_solver* solver = nullptr;
  int nn = 0, ng = 0, nnz = 0;
  int limit = 0;  // PARAMS.limit the twin was built with (selects the limiter function)
  std::vector<double> buf;
  ~Twin() {
    if (solver) fs_solver_destroy(solver);
    if (mesh) fs_mesh_destroy(mesh);
  }
};

struct GRID {
  int nn = 0, nn_ghost = 0, nedges = 0, nsubedges = 0, nbedges = 0;
  int iter = 1, nsteps = 0, citer = 0;
  int num_elem[kNumElemType] = {0, 0, 0};
  int** c2n = nullptr;   // c2n[Tri][t*3+k], c2n[Quad][t*4+k], t = 1..num_elem
  int nb = 0;
  int* nbs = nullptr;    // segments per boundary, b = 1..nb
  int*** bs = nullptr;   // bs[b][s][0..1], s = 1..nbs[b]
  int* bbc = nullptr;    // boundary condition per boundary, b = 1..nb
  double* x = nullptr;   // (nn+nn_ghost+1)*2
  double* Q = nullptr;
  double* nQ = nullptr;
  double* dQ = nullptr;
  double* R = nullptr;
  double* grad = nullptr;
  double* hess = nullptr;
  double* phi = nullptr;
  double* point_value = nullptr;
  double* ev_con = nullptr;
  double* dt = nullptr;
  double cfl = 0.;
  // fields of the per-face functions and the setup copies (1-based, slot 0
  // unused, as include/grid.h:224-306 lays them out)
  int* edges = nullptr;          // (nedges+1)*2
  int* subedges = nullptr;       // (nsubedges+1)*2: edge, element
  double* xm_subedges = nullptr; // (nsubedges+1)*2
  double* el_cent = nullptr;     // (elements+1)*2
  int* bedges = nullptr;         // (nbedges+1)*5: real node, ghost, element, boundary, segment
  double* xn_bedges = nullptr;   // (nbedges+1)*3: outward unit normal, length
  int* node_order = nullptr;     // nn+nn_ghost+1
  double* edge_weight = nullptr; // (nedges+1)*4
};

struct SYS_MEM {
  int* ia = nullptr;     // nn+2
  int* iau = nullptr;    // nn+1
  int* ja = nullptr;     // nnz, 1-based node ids
  double* LHS = nullptr; // nnz*16 (0-based slots)
  double* RHS = nullptr; // (nn+1)*4
};

namespace detail {

// device twins, keyed by the address of the caller's GRID
inline std::mutex& twin_mutex() {
  static std::mutex m;
  return m;
}
inline std::unordered_map<const void*, Twin*>& twins() {
  static std::unordered_map<const void*, Twin*> t;
  return t;
}
inline Twin* find_twin(const void* g) {
  std::lock_guard<std::mutex> lk(twin_mutex());
  auto it = twins().find(g);
  return it == twins().end() ? nullptr : it->second;
}

// the 24 PARAMS fields by name (include/params.h:51-74)
template <class P>
fs_params to_fs(const P& p) {
  fs_params q;
  q.nsteps = p.nsteps; q.iter = p.iter; q.jufreq = p.jufreq; q.CFL = p.CFL; q.ramp = p.ramp; q.gamma = p.gamma;
  q.alpha = p.alpha; q.mach = p.mach; q.subits = p.subits; q.unsteady = p.unsteady; q.tacc = p.tacc;
  q.mintime = p.mintime; q.method = p.method; q.order = p.order; q.foits = p.foits; q.soits = p.soits;
  q.toits = p.toits; q.isrestart = p.isrestart; q.limit = p.limit; q.reorder = p.reorder; q.ic = p.ic;
  q.p_or_cv = p.p_or_cv; q.start_post_pro = p.start_post_pro; q.post_pro_freq = p.post_pro_freq;
  return q;
}

[[noreturn]] inline void fatal(const char* where) {
  std::printf("FATAL ERROR: %s: %s\n", where, fs_last_error());
  std::fflush(stdout);
  std::exit(1);
}
inline void check(int rc, const char* where) {
  if (rc != FS_OK) fatal(where);
}
inline double* lazy(double*& p, size_t n) {
  if (!p) {
    p = static_cast<double*>(std::calloc(n, sizeof(double)));
    if (!p) {
      std::printf("MEMORY ERROR: COULD NOT ALLOCATE %zu doubles.\n", n);
      std::exit(1);
    }
  }
  return p;
}

template <class G>
Twin& attached(G* g, const char* where) {
  Twin* t = find_twin(g);
  if (!t) {
    std::printf("FATAL ERROR: %s: the GRID has no device twin yet (call solve or build_matrix first)\n", where);
    std::exit(1);
  }
  return *t;
}

template <class G>
Twin& twin(G* g, const fs_params& p) {
  if (Twin* have = find_twin(g)) return *have;
  std::vector<int> tri, quad, bseg, boff(1, 0), bbc;
  for (int t = 1; t <= g->num_elem[kTri]; ++t)
    for (int k = 0; k < 3; ++k) tri.push_back(g->c2n[kTri][t * 3 + k] - 1);
  for (int t = 1; t <= g->num_elem[kQuad]; ++t)
    for (int k = 0; k < 4; ++k) quad.push_back(g->c2n[kQuad][t * 4 + k] - 1);
  for (int b = 1; b <= g->nb; ++b) {
    for (int s = 1; s <= g->nbs[b]; ++s) {
      bseg.push_back(g->bs[b][s][0] - 1);
      bseg.push_back(g->bs[b][s][1] - 1);
    }
    boff.push_back((int)bseg.size() / 2);
    bbc.push_back(g->bbc[b]);
  }
  std::vector<double> x(2 * (size_t)g->nn);
  for (int n = 0; n < g->nn; ++n) {
    x[2 * n] = g->x[2 * (n + 1)];
    x[2 * n + 1] = g->x[2 * (n + 1) + 1];
  }
  Twin* t = new Twin();
  check(fs_mesh_create(g->nn, x.data(), g->num_elem[kTri], tri.data(), g->num_elem[kQuad], quad.data(), g->nb,
                       boff.data(), bseg.data(), bbc.data(), &t->mesh),
        "building the device mesh");
  fs_params q = p;
  q.reorder = 0;  // the GRID is already in its solver numbering
  check(fs_mesh_prepare(t->mesh, &q), "preparing the device mesh");
  fs_mesh_info mi;
  check(fs_mesh_get_info(t->mesh, &mi), "mesh info");
  check(fs_solver_create(t->mesh, &q, FS_PREC_F64, &t->solver), "creating the device solver");
  t->limit = p.limit;
  t->nn = mi.nn;
  t->ng = mi.ng;
  t->nnz = mi.nnz;
  g->nn_ghost = mi.ng;
  g->nedges = mi.ne;
  g->nsubedges = mi.nsub;
  g->nbedges = mi.nbe;
  std::lock_guard<std::mutex> lk(twin_mutex());
  twins()[g] = t;
  return *t;
}

// 1-based host array with `per` values per node (+ ghosts) <-> device field
inline void put(Twin& t, int field, const double* host, int per, bool ghosts) {
  size_t n = (size_t)per * (t.nn + (ghosts ? t.ng : 0));
  t.buf.assign(host + per, host + per + n);
  check(fs_solver_set(t.solver, field, t.buf.data(), (int64_t)n), "host-to-device copy");
}
inline void get(Twin& t, int field, double* host, int per, bool ghosts) {
  size_t n = (size_t)per * (t.nn + (ghosts ? t.ng : 0));
  t.buf.resize(n);
  check(fs_solver_get(t.solver, field, t.buf.data(), (int64_t)n), "device-to-host copy");
  std::memcpy(host + per, t.buf.data(), n * sizeof(double));
}
inline void put_phi(Twin& t, const double* phi) {
  t.buf.resize(t.nn);
  for (int n = 0; n < t.nn; ++n) t.buf[n] = phi[8 * (n + 1)];
  check(fs_solver_set(t.solver, FS_F_PHI, t.buf.data(), t.nn), "host-to-device copy");
}
inline void get_phi(Twin& t, double* phi) {
  t.buf.resize(t.nn);
  check(fs_solver_get(t.solver, FS_F_PHI, t.buf.data(), t.nn), "device-to-host copy");
  for (int n = 0; n < t.nn; ++n) phi[8 * (n + 1)] = phi[8 * (n + 1) + 1] = t.buf[n];
}
inline void put_hess(Twin& t, const double* hess) {
  t.buf.resize(20 * (size_t)t.nn);
  for (int n = 0; n < t.nn; ++n)
    for (int v = 0; v < 4; ++v)
      for (int r = 0; r < 5; ++r) t.buf[20 * (size_t)n + 5 * v + r] = hess[36 * (size_t)(n + 1) + 9 * v + r];
  check(fs_solver_set(t.solver, FS_F_HESS, t.buf.data(), 20 * (int64_t)t.nn), "host-to-device copy");
}
inline void get_hess(Twin& t, double* hess) {
  t.buf.resize(20 * (size_t)t.nn);
  check(fs_solver_get(t.solver, FS_F_HESS, t.buf.data(), 20 * (int64_t)t.nn), "device-to-host copy");
  for (int n = 0; n < t.nn; ++n)
    for (int v = 0; v < 4; ++v)
      for (int r = 0; r < 5; ++r) hess[36 * (size_t)(n + 1) + 9 * v + r] = t.buf[20 * (size_t)n + 5 * v + r];
}
// the system arrays the reference's build_matrix allocates on first use
// (init_memory.C:173-186): LHS of nnz 4x4 blocks, RHS of nn+1 4-vectors
template <class S>
void lazy_system(Twin& t, S* s) {
  if (!s) return;
  if (!s->LHS) lazy(s->LHS, 16 * (size_t)t.nnz);
  if (!s->RHS) lazy(s->RHS, 4 * ((size_t)t.nn + 1));
}

template <class S>
void put_lhs(Twin& t, const S* s) {
  if (s && s->LHS) check(fs_solver_set(t.solver, FS_F_LHS, s->LHS, 16 * (int64_t)t.nnz), "host-to-device copy");
}
template <class S>
void get_lhs(Twin& t, S* s) {
  if (s && s->LHS) check(fs_solver_get(t.solver, FS_F_LHS, s->LHS, 16 * (int64_t)t.nnz), "device-to-host copy");
}
template <class G>
void lazy_all(G* g) {
  size_t nn1 = (size_t)g->nn + 1;
  lazy(g->dQ, 4 * nn1);
  lazy(g->R, 4 * nn1);
  lazy(g->grad, 8 * nn1);
  lazy(g->phi, 8 * nn1);
  lazy(g->hess, 36 * nn1);
  lazy(g->point_value, 4 * nn1);
  lazy(g->ev_con, nn1);
  lazy(g->dt, nn1);
}

}  // namespace detail

// flux.C:299-541: host evaluation of the same function the device kernels use
inline int Roe_flux_centered(double nx, double ny, double len, double gamma, double qleft[4], double qright[4],
                             double flux[4]) {
  fsd::roe_flux(nx, ny, len, gamma, qleft, qright, flux);
  return 1;
}

// solve.C:44-933: one pseudo-time step on the device; Q, nQ, dQ, R and the
// reconstruction fields of the last Newton iteration come back to the GRID.
// Returns ||dQ||_2 of the last Newton iteration.
template <class G, class S, class P>
double solve(G* grid, S* smem, P p, int ts) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::lazy(grid->nQ, 4 * ((size_t)grid->nn + t.ng + 1));
  detail::put(t, FS_F_Q, grid->Q, 4, true);
  double dq = 0, r2 = 0;
  detail::check(fs_solve(t.solver, ts, &dq, &r2), "solve");
  detail::lazy_system(t, smem);
  detail::get(t, FS_F_Q, grid->Q, 4, true);
  detail::get(t, FS_F_NQ, grid->nQ, 4, true);
  detail::get(t, FS_F_DQ, grid->dQ, 4, false);
  detail::get(t, FS_F_R, grid->R, 4, false);
  detail::get(t, FS_F_GRAD, grid->grad, 8, false);
  detail::get_phi(t, grid->phi);
  detail::get_hess(t, grid->hess);
  detail::get(t, FS_F_PV, grid->point_value, 4, false);
  detail::get(t, FS_F_EVCON, grid->ev_con, 1, false);
  detail::get(t, FS_F_DT, grid->dt, 1, false);
  if (smem && smem->RHS) detail::get(t, FS_F_RHS, smem->RHS, 4, false);
  detail::get_lhs(t, smem);
  grid->citer = ts;
  return dq;
}

// init_memory.C:159-310: ev_con, dt and the V/dt diagonal at grid->citer
template <class G, class P, class S>
void build_matrix(G* grid, P p, S* smem) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::check(fs_build_matrix(t.solver, grid->citer), "build_matrix");
  detail::lazy_system(t, smem);
  detail::get(t, FS_F_EVCON, grid->ev_con, 1, false);
  detail::get(t, FS_F_DT, grid->dt, 1, false);
  detail::get_lhs(t, smem);
}

// gradient.C:363-403 over nQ
template <class G>
void Compute_Gradient_LeastSquares(G* grid) {
  Twin& t = detail::attached(grid, "Compute_Gradient_LeastSquares");
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::check(fs_gradient(t.solver), "Compute_Gradient_LeastSquares");
  detail::get(t, FS_F_GRAD, grid->grad, 8, false);
}

// compute_derivatives.C:1438-1534 (order_desired 3: the k=2 derivatives)
template <class G, class P>
void Compute_Derivatives(G* grid, P p, int order_desired) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  if (order_desired != 3) detail::check(FS_ERR_UNSUPPORTED, "Compute_Derivatives supports order_desired = 3");
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::check(fs_derivatives(t.solver), "Compute_Derivatives");
  detail::get_hess(t, grid->hess);
  detail::get(t, FS_F_PV, grid->point_value, 4, false);
}

// limiter.C:39-497 / limiter.C:511-983 at the points of the phase of grid->citer
template <class G, class P>
void compute_limiter_barth(G* grid, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  if (t.limit == 2) detail::check(FS_ERR_ARG, "compute_limiter_barth on a GRID attached with limit = 2");
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::put(t, FS_F_GRAD, grid->grad, 8, false);
  detail::put_hess(t, grid->hess);
  detail::put(t, FS_F_PV, grid->point_value, 4, false);
  detail::check(fs_limiter(t.solver, grid->citer), "compute_limiter_barth");
  detail::get_phi(t, grid->phi);
}
template <class G, class P>
void compute_limiter_venkatakrishnan(G* grid, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  if (t.limit != 2) detail::check(FS_ERR_ARG, "compute_limiter_venkatakrishnan needs a GRID attached with limit = 2");
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::put(t, FS_F_GRAD, grid->grad, 8, false);
  detail::put_hess(t, grid->hess);
  detail::put(t, FS_F_PV, grid->point_value, 4, false);
  detail::check(fs_limiter(t.solver, grid->citer), "compute_limiter_venkatakrishnan");
  detail::get_phi(t, grid->phi);
}

// residual.C:40-342: face fluxes into R (and the Jacobian into LHS)
template <class G, class S, class P>
void build_interior_residuals(G* grid, S* smem, P p, int jupdate, int ts) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::put(t, FS_F_GRAD, grid->grad, 8, false);
  detail::put_phi(t, grid->phi);
  if (jupdate) detail::put_lhs(t, smem);
  detail::check(fs_interior_residual(t.solver, jupdate, ts), "build_interior_residuals");
  detail::get(t, FS_F_R, grid->R, 4, false);
  if (jupdate) detail::get_lhs(t, smem);
}

// residual.C:359-1230: ghost update with memory, boundary fluxes (and blocks)
template <class G, class S, class P>
void build_boundary_residuals(G* grid, S* smem, P p, int jupdate) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::put(t, FS_F_R, grid->R, 4, false);
  if (jupdate) detail::put_lhs(t, smem);
  detail::check(fs_boundary_residual(t.solver, jupdate), "build_boundary_residuals");
  detail::get(t, FS_F_R, grid->R, 4, false);
  detail::get(t, FS_F_NQ, grid->nQ, 4, true);
  if (jupdate) detail::get_lhs(t, smem);
}

// linear_algebra.C:122-401: the multicolor form of the symmetric block
// Gauss-Seidel sweeps on smem's system; dQ comes back to the GRID
template <class G, class S, class P>
double linear_solve(G* grid, S* smem, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::put_lhs(t, smem);
  if (smem && smem->RHS) detail::put(t, FS_F_RHS, smem->RHS, 4, false);
  double rms = 0;
  detail::check(fs_linear_solve(t.solver, &rms), "linear_solve");
  detail::get(t, FS_F_DQ, grid->dQ, 4, false);
  return rms;
}

// grid_io.C:1390-1697: lift (Forces[0]) and drag (Forces[1]) over the walls
// (odd bc; straight and curved) from the GRID's state after step grid->citer
template <class G, class P>
void Compute_Body_Forces(G* grid, P p, double* Forces) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy_all(grid);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::put(t, FS_F_GRAD, grid->grad, 8, false);
  detail::put_phi(t, grid->phi);
  detail::put_hess(t, grid->hess);
  detail::put(t, FS_F_PV, grid->point_value, 4, false);
  detail::check(fs_body_forces(t.solver, grid->citer, Forces), "Compute_Body_Forces");
}

// grid_io.C:841-1375: surface_<b>_pressure.dat for every wall in the working
// directory, Cp of the GRID's Q at the unique segment nodes
template <class G, class P>
void write_surface_cp(G* grid, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::put(t, FS_F_Q, grid->Q, 4, true);
  detail::check(fs_write_surface_cp(t.solver, "."), "write_surface_cp");
}

// include/curved_boundaries.h:25-33 (boundary codes 11-22): the analytic
// curved walls on the host; the GRID argument is unused, as in the reference
template <class G>
void curved_boundary_midpoint(G*, int bc, double* xL, double* xR, double* xM) {
  detail::check(fs_curved_midpoint(bc, xL, xR, xM), "curved_boundary_midpoint");
}
template <class G>
void curved_boundary_gauss_points(G*, int bc, double* xL, double* xR, double* GP) {
  detail::check(fs_curved_gauss_points(bc, xL, xR, GP), "curved_boundary_gauss_points");
}
template <class G>
void curved_boundary_normal_vector(G*, int bc, double* x, double* nx, double* ny) {
  detail::check(fs_curved_normal(bc, x, nx, ny), "curved_boundary_normal_vector");
}
template <class G>
void curved_boundary_arclength(G*, int bc, double* xL, double* xR, double* arclength) {
  detail::check(fs_curved_arclength(bc, xL, xR, arclength), "curved_boundary_arclength");
}

namespace detail {
[[noreturn]] inline void unsupported(const char* where, const char* what) {
  std::printf("FATAL ERROR: %s: %s\n", where, what);
  std::fflush(stdout);
  std::exit(1);
}
// Reconstruction order of `node` at pseudo-time step `ts` (reconstruct.C:72-
// 171, 1161-1188): 0 nodal, 1 linear, 3 quadratic; -1 for order 0 (no
// reconstruction).
template <class G, class P>
int recon_order(const G* g, const P& p, int ts, int node, const char* where) {
  if (p.order == 0) return -1;
  if (p.order == 1) return ts <= p.foits ? 0 : 1;
  if (p.order != 3) unsupported(where, "only orders 0, 1 and 3 are on this path");
  if (ts <= p.foits) return 0;
  if (ts <= p.foits + p.soits) return 1;
  const int r = g->node_order[node];
  if (r != 1 && r != 3) unsupported(where, "node order must be 1 or 3");
  return r;
}
// Q of `node` extrapolated by `dx` at order r (reconstruct.C:181-205,
// 222-254; PV_EXT: the quadratic starts from the point value)
template <class G>
void extrapolate(const G* g, int node, int r, const double* dx, double* Q) {
  if (r == 1) {
    fsd::extrapolate_linear(Q, g->phi[node * 8], g->grad + 8 * (size_t)node, dx[0], dx[1]);
  } else if (r == 3) {
    fsd::extrapolate_quadratic(Q, g->point_value + 4 * (size_t)node, g->phi[node * 8], g->phi[node * 8 + 1],
                               g->hess + 36 * (size_t)node, 9, dx[0], dx[1]);
  }
}
// slot of block (row, col) in the reference's CRS (init_memory.C:36-138);
// -1 when absent (the reference's search then writes nothing)
template <class S>
int block_slot(const S* s, int row, int col) {
  for (int i = s->ia[row]; i < s->ia[row + 1]; i++)
    if (s->ja[i] == col) return i;
  return -1;
}
// the reference's four block updates of one interior face (jacobian.C:
// 135-182): diag L += dL, diag R -= dR, (L,R) += dR, (R,L) -= dL
template <class S>
void scatter_blocks(S* s, int nodeL, int nodeR, const double* dL, const double* dR) {
  double* LHS = s->LHS;
  for (int k = 0; k < 16; k++) LHS[(size_t)s->iau[nodeL] * 16 + k] += dL[k];
  for (int k = 0; k < 16; k++) LHS[(size_t)s->iau[nodeR] * 16 + k] -= dR[k];
  for (int i = s->ia[nodeL]; i < s->ia[nodeL + 1]; i++)
    if (s->ja[i] == nodeR)
      for (int k = 0; k < 16; k++) LHS[(size_t)i * 16 + k] += dR[k];
  for (int i = s->ia[nodeR]; i < s->ia[nodeR + 1]; i++)
    if (s->ja[i] == nodeL)
      for (int k = 0; k < 16; k++) LHS[(size_t)i * 16 + k] -= dL[k];
}
inline void warn_nonfinite(const double* d, const char* which, int nodeL, int nodeR) {
  for (int k = 0; k < 16; k++)
    if (!std::isfinite(d[k]))
      std::printf("Problem DETECTED in %s[%d][%d]=%f  between nodeL=%d and nodeR=%d.\n", which, k / 4, k % 4,
                  (float)d[k], nodeL, nodeR);
}
}  // namespace detail

// reconstruct.C:38-446: limited face states of subedge `isedge` (1-based)
// at its face point for the phase of step ts; QL/QR start from nQ
template <class G, class P>
void Reconstruct(int isedge, int nodeL, int nodeR, double QL[4], double QR[4], G* grid, P p, int ts) {
  for (int j = 0; j < 4; j++) {
    QL[j] = grid->nQ[nodeL * 4 + j];
    QR[j] = grid->nQ[nodeR * 4 + j];
  }
  const int rL = detail::recon_order(grid, p, ts, nodeL, "Reconstruct");
  if (rL < 0) return;
  const int rR = detail::recon_order(grid, p, ts, nodeR, "Reconstruct");
  // face point: midway between the element centroid and the edge midpoint
  const int gelem = grid->subedges[isedge * 2 + 1];
  const double fc0 = (grid->el_cent[gelem * 2 + 0] + grid->xm_subedges[isedge * 2 + 0]) * 0.5;
  const double fc1 = (grid->el_cent[gelem * 2 + 1] + grid->xm_subedges[isedge * 2 + 1]) * 0.5;
  double dx[2] = {fc0 - grid->x[nodeL * 2 + 0], fc1 - grid->x[nodeL * 2 + 1]};
  detail::extrapolate(grid, nodeL, rL, dx, QL);
  dx[0] = fc0 - grid->x[nodeR * 2 + 0];
  dx[1] = fc1 - grid->x[nodeR * 2 + 1];
  detail::extrapolate(grid, nodeR, rR, dx, QR);
}

// reconstruct.C:1109-1399: limited states of both ends at the point GP, in
// the phase of grid->citer (as the reference reads it; ts is unused there
// too)
template <class G, class P>
void Reconstruct_Gauss(int isedge, int nodeL, int nodeR, double GP[2], double QL[4], double QR[4], G* grid, P p,
                       int ts) {
  (void)isedge;
  (void)ts;
  for (int j = 0; j < 4; j++) {
    QL[j] = grid->nQ[nodeL * 4 + j];
    QR[j] = grid->nQ[nodeR * 4 + j];
  }
  const int rL = detail::recon_order(grid, p, grid->citer, nodeL, "Reconstruct_Gauss");
  if (rL <= 0) return;
  const int rR = detail::recon_order(grid, p, grid->citer, nodeR, "Reconstruct_Gauss");
  double dx[2] = {GP[0] - grid->x[nodeL * 2 + 0], GP[1] - grid->x[nodeL * 2 + 1]};
  detail::extrapolate(grid, nodeL, rL, dx, QL);
  dx[0] = GP[0] - grid->x[nodeR * 2 + 0];
  dx[1] = GP[1] - grid->x[nodeR * 2 + 1];
  detail::extrapolate(grid, nodeR, rR, dx, QR);
}

// reconstruct.C:1708-1890: the limited state of a boundary node at GP in the
// phase of grid->citer
template <class G, class P>
void Reconstruct_Gauss_Boundary(int nodeL, double GP[2], double QL[4], G* grid, P p) {
  for (int j = 0; j < 4; j++) QL[j] = grid->nQ[nodeL * 4 + j];
  const int r = detail::recon_order(grid, p, grid->citer, nodeL, "Reconstruct_Gauss_Boundary");
  if (r <= 0) return;
  const double dx[2] = {GP[0] - grid->x[nodeL * 2 + 0], GP[1] - grid->x[nodeL * 2 + 1]};
  detail::extrapolate(grid, nodeL, r, dx, QL);
}

// jacobian.C:45-185: the forward-difference flux Jacobians of one interior
// face (flux = Roe_flux_centered(QL, QR)) added into smem->LHS
template <class G, class S, class P>
void compute_numerical_jacobian_roe(G* grid, S* smem, P p, int nodeL, int nodeR, double QL[4], double QR[4],
                                    double flux[4], double nx, double ny, double len) {
  (void)grid;
  double dL[16], dR[16];
  fsd::fd_jacobian(nx, ny, len, p.gamma, QL, QR, flux, dL, dR);
  detail::warn_nonfinite(dL, "dfdqL", nodeL, nodeR);
  detail::warn_nonfinite(dR, "dfdqR", nodeL, nodeR);
  detail::scatter_blocks(smem, nodeL, nodeR, dL, dR);
}

// jacobian.C:301-658: the analytic approximate Jacobians of one interior
// face, 0.5 (A(Q) +- |A_roe|) len, added into smem->LHS (flux is unused)
template <class G, class S, class P>
void compute_approximate_jacobian_roe(G* grid, S* smem, P p, int nodeL, int nodeR, double QL[4], double QR[4],
                                      double flux[4], double nx, double ny, double len) {
  (void)grid;
  (void)flux;
  double dL[16], dR[16];
  fsd::analytic_jacobian(nx, ny, len, p.gamma, QL, QR, dL, dR);
  detail::warn_nonfinite(dL, "dfdqL", nodeL, nodeR);
  detail::warn_nonfinite(dR, "dfdqR", nodeL, nodeR);
  detail::scatter_blocks(smem, nodeL, nodeR, dL, dR);
}

// jacobian.C:681-962: 0.5 (A(QL) + |A_roe|) len of boundary edge `ibedge`
// (1-based) added to its real node's diagonal block (QR: the ghost state)
template <class G, class S, class P>
void compute_boundary_approximate_jacobian_roe(G* grid, S* smem, P p, int ibedge, double QL[4], double QR[4],
                                               double flux[4], double nx, double ny, double len) {
  (void)flux;
  double d[16];
  fsd::boundary_jacobian(nx, ny, len, p.gamma, QL, QR, d);
  const int n = grid->bedges[ibedge * 5 + 0];
  for (int k = 0; k < 16; k++)
    if (!std::isfinite(d[k]))
      std::printf("Problem DETECTED in dfdqL[%d][%d]=%f  on bedge=%d - node = %d\n", k / 4, k % 4, (float)d[k],
                  ibedge, n);
  for (int k = 0; k < 16; k++) smem->LHS[(size_t)smem->iau[n] * 16 + k] += d[k];
}

// cvbc.C:502-891: the ghost state Qb of boundary edge `isubedge` (1-based)
// for the interior state Qi, with the edge's ghost in grid->nQ as the
// previous boundary state (its memory term)
template <class G, class P>
void Get_Boundary_Value(G* grid, P p, int isubedge, double Qi[4], double Qb[4]) {
  const int ghost = grid->bedges[isubedge * 5 + 1];
  const int b = grid->bedges[isubedge * 5 + 3];
  const int bc = grid->bbc[b];
  const double nx = grid->xn_bedges[isubedge * 3 + 0], ny = grid->xn_bedges[isubedge * 3 + 1];
  const double angle = p.alpha * M_PI / 180.;
  const double uinf = p.mach * cos(angle), vinf = p.mach * sin(angle), pinf = 1. / p.gamma;
  if (bc < 0 && bc % 2 != 0) detail::unsupported("Get_Boundary_Value", "an invalid boundary condition was set");
  for (int j = 0; j < 4; j++) Qb[j] = grid->nQ[((size_t)grid->nn + ghost) * 4 + j];
  if (!fsd::ghost_state(bc, nx, ny, p.gamma, uinf, vinf, pinf, Qi, Qb))
    detail::unsupported("Get_Boundary_Value", "the eigenvalues do not correspond to a valid inflow/outflow state");
}

// jacobian.C:205-277: forward differences (eps 1e-8) of the boundary flux
// Roe_flux_centered(QL, Get_Boundary_Value(QL)) added to the real node's
// diagonal block; flux is the unperturbed boundary flux
template <class G, class S, class P>
void compute_boundary_numerical_jacobian_roe(G* grid, S* smem, P p, int ibedge, double QL[4], double flux[4],
                                             double nx, double ny, double len) {
  const double eps = 1.0e-08;
  const int n = grid->bedges[ibedge * 5 + 0];
  double d[16], Qp[4], Qb[4], fp[4];
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) Qp[j] = QL[j];
    Qp[i] += eps;
    Get_Boundary_Value(grid, p, ibedge, Qp, Qb);
    fsd::roe_flux(nx, ny, len, p.gamma, Qp, Qb, fp);
    for (int j = 0; j < 4; j++) d[4 * j + i] = (fp[j] - flux[j]) / eps;
  }
  for (int k = 0; k < 16; k++)
    if (!std::isfinite(d[k]))
      std::printf("Problem DETECTED in dfdqL[%d][%d]=%f on bedge %d, real node = %d.\n", k / 4, k % 4, (float)d[k],
                  ibedge, n);
  for (int k = 0; k < 16; k++) smem->LHS[(size_t)smem->iau[n] * 16 + k] += d[k];
}

// cvbc.C:37-481: the ghost states of nQ (with their memory) on the device
template <class G, class P>
void Set_Boundary_Values(G* grid, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::check(fs_set_boundary_values(t.solver), "Set_Boundary_Values");
  detail::get(t, FS_F_NQ, grid->nQ, 4, true);
}

// cv_calc.C:790-965: ev_con of nQ on the device
template <class G, class P>
void cv_calc_eigenvalue_contribution(G* grid, P p) {
  Twin& t = detail::twin(grid, detail::to_fs(p));
  detail::lazy(grid->ev_con, (size_t)grid->nn + 1);
  detail::put(t, FS_F_NQ, grid->nQ, 4, true);
  detail::check(fs_eigenvalue_contribution(t.solver), "cv_calc_eigenvalue_contribution");
  detail::get(t, FS_F_EVCON, grid->ev_con, 1, false);
}

namespace detail {
// a prepared host mesh of the GRID in its current numbering: the GRID's
// twin when it has one, else a temporary (released by the caller)
template <class G>
fs_mesh* prepared_mesh(G* g, const fs_params& p, bool* temporary) {
  if (Twin* t = find_twin(g)) {
    *temporary = false;
    return t->mesh;
  }
  std::vector<int> tri, quad, bseg, boff(1, 0), bbc;
  for (int e = 1; e <= g->num_elem[kTri]; ++e)
    for (int k = 0; k < 3; ++k) tri.push_back(g->c2n[kTri][e * 3 + k] - 1);
  for (int e = 1; e <= g->num_elem[kQuad]; ++e)
    for (int k = 0; k < 4; ++k) quad.push_back(g->c2n[kQuad][e * 4 + k] - 1);
  for (int b = 1; b <= g->nb; ++b) {
    for (int s = 1; s <= g->nbs[b]; ++s) {
      bseg.push_back(g->bs[b][s][0] - 1);
      bseg.push_back(g->bs[b][s][1] - 1);
    }
    boff.push_back((int)bseg.size() / 2);
    bbc.push_back(g->bbc[b]);
  }
  std::vector<double> x(2 * (size_t)g->nn);
  for (int n = 0; n < g->nn; ++n) {
    x[2 * n] = g->x[2 * (n + 1)];
    x[2 * n + 1] = g->x[2 * (n + 1) + 1];
  }
  fs_mesh* m = nullptr;
  check(fs_mesh_create(g->nn, x.data(), g->num_elem[kTri], tri.data(), g->num_elem[kQuad], quad.data(), g->nb,
                       boff.data(), bseg.data(), bbc.data(), &m),
        "building the host mesh");
  fs_params q = p;
  q.reorder = 0;
  check(fs_mesh_prepare(m, &q), "preparing the host mesh");
  *temporary = true;
  return m;
}
inline const void* mesh_array(fs_mesh* m, int which, int64_t* count) {
  const void* d = nullptr;
  int dtype = 0;
  check(fs_mesh_get_array(m, which, &d, count, &dtype), "mesh array");
  return d;
}
}  // namespace detail

// gradient.C:923-1121: the LSQ edge weights (4 per edge) of the GRID's
// edges, from the host preparation (the same closed form, bit for bit)
template <class G>
void Compute_Edge_Weights(G* grid) {
  fs_params p;
  fs_params_default(&p);
  bool temporary = false;
  fs_mesh* m = detail::prepared_mesh(grid, p, &temporary);
  int64_t cnt = 0;
  const double* w = static_cast<const double*>(detail::mesh_array(m, FS_M_EDGE_WEIGHT, &cnt));
  detail::lazy(grid->edge_weight, (size_t)cnt + 4);
  std::memcpy(grid->edge_weight + 4, w, (size_t)cnt * sizeof(double));
  if (temporary) fs_mesh_destroy(m);
}

// init_memory.C:36-138: the block-CSR pattern (ia, iau, ja; the diagonal
// last in each row), allocated with malloc where null as the reference does
template <class G, class P, class S>
void Build_CRS(G* grid, P p, S* smem) {
  bool temporary = false;
  fs_mesh* m = detail::prepared_mesh(grid, detail::to_fs(p), &temporary);
  int64_t nia = 0, niau = 0, nja = 0;
  const int* ia = static_cast<const int*>(detail::mesh_array(m, FS_M_IA, &nia));
  const int* iau = static_cast<const int*>(detail::mesh_array(m, FS_M_IAU, &niau));
  const int* ja = static_cast<const int*>(detail::mesh_array(m, FS_M_JA, &nja));
  const int nn = grid->nn;
  if (!smem->ia) smem->ia = static_cast<int*>(std::malloc((size_t)(nn + 2) * sizeof(int)));
  if (!smem->iau) smem->iau = static_cast<int*>(std::malloc((size_t)(nn + 1) * sizeof(int)));
  if (!smem->ja) smem->ja = static_cast<int*>(std::malloc((size_t)nja * sizeof(int)));
  if (!smem->ia || !smem->iau || !smem->ja) detail::unsupported("Build_CRS", "MEMORY ERROR: COULD NOT ALLOCATE");
  // 1-based rows, 0-based slots, 1-based column nodes
  smem->ia[0] = 0;
  for (int n = 1; n <= nn + 1; n++) smem->ia[n] = ia[n - 1];
  smem->iau[0] = 0;
  for (int n = 1; n <= nn; n++) smem->iau[n] = iau[n - 1];
  for (int64_t k = 0; k < nja; k++) smem->ja[k] = ja[k] + 1;
  if (temporary) fs_mesh_destroy(m);
}

// releases the device twin of a GRID
template <class G>
void detach(G* grid) {
  Twin* t = nullptr;
  {
    std::lock_guard<std::mutex> lk(detail::twin_mutex());
    auto it = detail::twins().find(grid);
    if (it == detail::twins().end()) return;
    t = it->second;
    detail::twins().erase(it);
  }
  delete t;
}

}  // namespace fsref

#endif
