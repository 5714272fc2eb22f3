// C-ABI implementation: mesh handles (host setup) and the single-mesh solver
// (a one-instance device engine with its own HIP stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/fs_curved.hpp"
#include "../host/fs_setup.hpp"
#include "../kernels/fs_kernels.hpp"
#include "fs_capi.h"
#include "fs_capi_internal.hpp"
#include "fs_engine.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace fse {
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

const char* last_error_text() { return g_err.c_str(); }

const char* device_error_message(int e) {
  if (e & 1) return "boundary state outside the characteristic cases (cvbc.C:330-382 exit)";
  if (e & 4) return "a pre-multiplied sweep operand outside the FP16 range (FS_PREC_MIXED16)";
  return "non-positive or non-finite time step (init_memory.C:257-268 exit)";
}
}  // namespace fse

#define HIP_TRY FSE_HIP

using fse::from_params;
using fse::set_error;
using fse::to_params;

namespace {
int fail(int code, const std::string& msg) { return set_error(code, msg); }
}  // namespace

struct fs_solver {
  fse::Engine eng;
  bool graph = false;
  bool timing = false;
};

namespace {

int need_f64(const fs_solver* s) {
  if (s->eng.precision != FS_PREC_F64)
    return fail(FS_ERR_ARG, "the FP64 Jacobian functions need an FP64 solver (mixed precision keeps FP32 blocks)");
  return FS_OK;
}

// controls for a single-function mirror call at pseudo-time step ts
fsk::InstParams mirror_params(fs_solver* s, int ts) {
  s->eng.inst[0].ts = ts;
  s->eng.inst[0].active = true;
  return s->eng.params_for(0, 1);
}

int finish(fs_solver* s, bool errors) {
  HIP_TRY(hipGetLastError());
  if (!errors) return s->eng.sync();
  int rc = s->eng.fetch_errors();
  if (rc) return rc;
  int e = s->eng.inst[0].err;
  if (e) return fail(FS_ERR_STATE, fse::device_error_message(e));
  return FS_OK;
}

}  // namespace

extern "C" {

const char* fs_last_error(void) { return g_err.c_str(); }

int fs_hbm_probe(int64_t bytes, int reps, double* read_gbs, double* copy_gbs) {
  if (bytes < (1 << 20) || reps < 1 || !read_gbs || !copy_gbs) return fail(FS_ERR_ARG, "bad arguments");
  const size_t n16 = (size_t)bytes / 16, nb = n16 * 16;
  void *a = nullptr, *b = nullptr;
  float* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FS_OK;
  auto release = [&] {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(out);
  };
  if (hipMalloc(&a, nb) != hipSuccess || hipMalloc(&b, nb) != hipSuccess || hipMalloc((void**)&out, 64) != hipSuccess ||
      hipMemset(a, 0, nb) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    release();
    return fail(FS_ERR_HIP, "probe buffers: " + std::string(hipGetErrorString(hipGetLastError())));
  }
  float ms_r = 0.f, ms_c = 0.f;
  fsk::launch_probe_read(a, n16, out, nullptr);   // warm-up
  fsk::launch_probe_copy(a, b, n16, nullptr);
  if (hipEventRecord(e0, nullptr) != hipSuccess) rc = FS_ERR_HIP;
  for (int k = 0; k < reps && !rc; ++k) fsk::launch_probe_read(a, n16, out, nullptr);
  if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms_r, e0, e1) != hipSuccess))
    rc = FS_ERR_HIP;
  if (!rc && hipEventRecord(e0, nullptr) != hipSuccess) rc = FS_ERR_HIP;
  for (int k = 0; k < reps && !rc; ++k) fsk::launch_probe_copy(a, b, n16, nullptr);
  if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms_c, e0, e1) != hipSuccess))
    rc = FS_ERR_HIP;
  release();
  if (rc) return fail(rc, "probe launch failed");
  *read_gbs = (double)reps * nb / (ms_r * 1e-3) / 1e9;
  *copy_gbs = 2.0 * reps * nb / (ms_c * 1e-3) / 1e9;
  return FS_OK;
}

int fs_device_count(int* count) {
  if (!count) return fail(FS_ERR_ARG, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FS_OK;
}

void fs_params_default(fs_params* p) {
  if (p) from_params(fs::Params(), p);
}

int fs_params_read(const char* path, fs_params* p) {
  if (!path || !p) return fail(FS_ERR_ARG, "null argument");
  fs::Params q;
  std::string err;
  if (!fs::read_params(path, &q, &err)) return fail(FS_ERR_IO, err);
  from_params(q, p);
  return FS_OK;
}

int fs_params_write(const char* path, const fs_params* p) {
  if (!path || !p) return fail(FS_ERR_ARG, "null argument");
  if (!fs::write_params(path, to_params(p))) return fail(FS_ERR_IO, std::string("cannot write ") + path);
  return FS_OK;
}

int fs_mesh_read(const char* grid_path, const char* bc_path, fs_mesh** out) {
  if (!grid_path || !out) return fail(FS_ERR_ARG, "null argument");
  auto h = std::make_unique<fs_mesh>();
  std::string err;
  if (!fs::read_grid(grid_path, &h->m, &err)) return fail(FS_ERR_IO, err);
  if (bc_path && !fs::read_bc(bc_path, &h->m, &err)) return fail(FS_ERR_IO, err);
  *out = h.release();
  return FS_OK;
}

int fs_mesh_write(const fs_mesh* m, const char* grid_path, const char* bc_path) {
  if (!m || !grid_path) return fail(FS_ERR_ARG, "null argument");
  if (m->prepared) return fail(FS_ERR_ARG, "write the mesh before fs_mesh_prepare (preparation renumbers nodes)");
  if (!fs::write_grid(grid_path, m->m)) return fail(FS_ERR_IO, std::string("cannot write ") + grid_path);
  if (bc_path && !fs::write_bc(bc_path, m->m)) return fail(FS_ERR_IO, std::string("cannot write ") + bc_path);
  return FS_OK;
}

int fs_mesh_create_bump(int nx, int ny, double jitter, uint64_t seed, fs_mesh** out) {
  if (!out || nx < 3 || ny < 3 || jitter < 0 || jitter >= 0.5) return fail(FS_ERR_ARG, "bad bump-channel size/jitter");
  auto h = std::make_unique<fs_mesh>();
  fs::make_bump_channel(nx, ny, jitter, seed, &h->m);
  *out = h.release();
  return FS_OK;
}

int fs_mesh_create(int nn, const double* x, int ntri, const int* tri, int nquad, const int* quad, int nb,
                   const int* bseg_off, const int* bseg, const int* bbc, fs_mesh** out) {
  if (!out || nn < 3 || !x || ntri < 0 || nquad < 0 || ntri + nquad < 1 || nb < 1 || !bseg_off || !bbc ||
      (ntri > 0 && !tri) || (nquad > 0 && !quad))
    return fail(FS_ERR_ARG, "bad mesh arrays");
  auto h = std::make_unique<fs_mesh>();
  fs::Mesh& m = h->m;
  m.nn = nn;
  m.x.assign(x, x + 2 * (size_t)nn);
  m.ntri = ntri;
  m.tri.assign(tri, tri + 3 * (size_t)ntri);
  m.nquad = nquad;
  if (nquad > 0) m.quad.assign(quad, quad + 4 * (size_t)nquad);
  m.nb = nb;
  m.bseg_off.assign(bseg_off, bseg_off + nb + 1);
  if (m.bseg_off[0] != 0) return fail(FS_ERR_ARG, "bseg_off[0] must be 0");
  for (int b = 0; b < nb; ++b)
    if (m.bseg_off[b + 1] < m.bseg_off[b]) return fail(FS_ERR_ARG, "bseg_off not ascending");
  if (m.bseg_off[nb] > 0 && !bseg) return fail(FS_ERR_ARG, "null bseg");
  m.bseg.assign(bseg, bseg + 2 * (size_t)m.bseg_off[nb]);
  m.bbc.assign(bbc, bbc + nb);
  for (int v : m.tri) if (v < 0 || v >= nn) return fail(FS_ERR_ARG, "triangle node out of range");
  for (int v : m.quad) if (v < 0 || v >= nn) return fail(FS_ERR_ARG, "quad node out of range");
  for (int v : m.bseg) if (v < 0 || v >= nn) return fail(FS_ERR_ARG, "segment node out of range");
  *out = h.release();
  return FS_OK;
}

int fs_mesh_create_airfoil(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed, fs_mesh** out) {
  if (!out || nwall < 8 || nrad < 3 || !(h0 > 0) || !(rmax > 1) || jitter < 0 || jitter >= 0.5)
    return fail(FS_ERR_ARG, "bad airfoil O-mesh arguments");
  auto h = std::make_unique<fs_mesh>();
  fs::make_airfoil_omesh(nwall, nrad, h0, rmax, jitter, seed, &h->m);
  *out = h.release();
  return FS_OK;
}

int fs_mesh_create_airfoil_split(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed,
                                 int upper_bc, int lower_bc, fs_mesh** out) {
  if (!out || nwall < 8 || nwall % 2 || nrad < 3 || !(h0 > 0) || !(rmax > 1) || jitter < 0 || jitter >= 0.5)
    return fail(FS_ERR_ARG, "bad airfoil O-mesh arguments");
  auto wall_ok = [](int bc, bool upper) { return bc == 1 || bc == (upper ? 17 : 19); };
  if (!wall_ok(upper_bc, true) || !wall_ok(lower_bc, false))
    return fail(FS_ERR_ARG, "upper/lower wall codes must be 1 (polyline) or 17/19 (NACA 0012 upper/lower)");
  auto h = std::make_unique<fs_mesh>();
  fs::make_airfoil_omesh_split(nwall, nrad, h0, rmax, jitter, seed, upper_bc, lower_bc, &h->m);
  *out = h.release();
  return FS_OK;
}

int fs_mesh_create_annulus(int ntheta, int nrad, double r_in, double r_out, double jitter, uint64_t seed, int bc_in,
                           int bc_out, fs_mesh** out) {
  if (!out || ntheta < 8 || ntheta % 2 || nrad < 2 || !(r_in > 0) || !(r_out > r_in) || jitter < 0 ||
      jitter >= 0.5)
    return fail(FS_ERR_ARG, "bad annulus arguments");
  // Codes 13-16 are left out: their arclength takes the angle difference
  // without the branch-cut handling of codes 11/12 and 21/22
  // (curved_boundaries.C:1063-1093), so a closed ring gets one segment about
  // 2 pi R long next to the angle pi.
  auto ring_ok = [](int bc, double r) {
    if (bc <= 10) return bc >= 0;
    const double R = (bc == 11 || bc == 12) ? 1.0 : (bc == 21 || bc == 22) ? 300.0 : -1.0;
    return R == r;
  };
  if (!ring_ok(bc_in, r_in) || !ring_ok(bc_out, r_out))
    return fail(FS_ERR_ARG, "ring codes: 0-10, or 11/12 at radius 1, or 21/22 at radius 300");
  auto h = std::make_unique<fs_mesh>();
  fs::make_annulus(ntheta, nrad, r_in, r_out, jitter, seed, bc_in, bc_out, &h->m);
  *out = h.release();
  return FS_OK;
}

int fs_curved_midpoint(int bc, const double* xL, const double* xR, double* xM) {
  if (!xL || !xR || !xM) return fail(FS_ERR_ARG, "null argument");
  if (!fs::curved_bc_known(bc)) return fail(FS_ERR_ARG, "not a curved boundary code (11-22)");
  fs::curved_midpoint(bc, xL, xR, xM);
  return FS_OK;
}
int fs_curved_gauss_points(int bc, const double* xL, const double* xR, double* GP) {
  if (!xL || !xR || !GP) return fail(FS_ERR_ARG, "null argument");
  if (!fs::curved_bc_known(bc)) return fail(FS_ERR_ARG, "not a curved boundary code (11-22)");
  std::string err;
  if (!fs::curved_gauss_points(bc, xL, xR, GP, &err)) return fail(FS_ERR_MESH, err);
  return FS_OK;
}
int fs_curved_normal(int bc, const double* x, double* nx, double* ny) {
  if (!x || !nx || !ny) return fail(FS_ERR_ARG, "null argument");
  if (!fs::curved_bc_known(bc)) return fail(FS_ERR_ARG, "not a curved boundary code (11-22)");
  fs::curved_normal(bc, x, nx, ny);
  return FS_OK;
}
int fs_curved_arclength(int bc, const double* xL, const double* xR, double* len) {
  if (!xL || !xR || !len) return fail(FS_ERR_ARG, "null argument");
  if (!fs::curved_bc_known(bc)) return fail(FS_ERR_ARG, "not a curved boundary code (11-22)");
  *len = fs::curved_arclength(bc, xL, xR);
  return FS_OK;
}

int fs_mesh_prepare(fs_mesh* m, const fs_params* p) {
  if (!m || !p) return fail(FS_ERR_ARG, "null argument");
  if (m->prepared) return fail(FS_ERR_ARG, "mesh already prepared");
  std::string err;
  if (!fs::prepare(&m->m, to_params(p), &err)) return fail(FS_ERR_MESH, err);
  m->prepared = true;
  m->order_prepared = p->order;
  return FS_OK;
}

int fs_restart_write(const fs_mesh* m, const char* grid_path, const double* Q, int64_t n, int citer, double cfl,
                     int with_ghosts) {
  if (!m || !grid_path || !Q) return fail(FS_ERR_ARG, "null argument");
  if (!m->prepared) return fail(FS_ERR_ARG, "restart files are written from a prepared mesh (solver numbering)");
  const fs::Mesh& g = m->m;
  if (n != 4 * ((int64_t)g.nn + g.ng)) return fail(FS_ERR_ARG, "size mismatch: expected 4*(nn+ng)");
  if (!fs::write_restart(grid_path, g, Q, g.ng, citer, cfl, with_ghosts != 0))
    return fail(FS_ERR_IO, std::string("cannot write ") + grid_path);
  return FS_OK;
}

int fs_mesh_restart_info(const fs_mesh* m, int* has_restart, int* citer, double* cfl, int* has_ghosts) {
  if (!m) return fail(FS_ERR_ARG, "null mesh");
  const fs::Mesh& g = m->m;
  if (has_restart) *has_restart = g.has_restart ? 1 : 0;
  if (citer) *citer = g.restart_citer;
  if (cfl) *cfl = g.restart_cfl;
  if (has_ghosts) *has_ghosts = g.restart_ghost_Q.empty() ? 0 : 1;
  return FS_OK;
}

int fs_mesh_get_info(const fs_mesh* m, fs_mesh_info* info) {
  if (!m || !info) return fail(FS_ERR_ARG, "null argument");
  const fs::Mesh& g = m->m;
  info->nn = g.nn; info->ng = g.ng; info->ne = g.ne; info->nsub = g.nsub; info->nbe = g.nbe;
  info->nnz = (int)g.ja.size(); info->ntri = g.ntri; info->nquad = g.nquad; info->nb = g.nb;
  info->ncolor = g.ncolor; info->nstencil = (int)g.st_nodes.size();
  info->prepared = m->prepared ? 1 : 0;
  info->order_prepared = m->order_prepared;
  return FS_OK;
}

int fs_mesh_get_array(const fs_mesh* m, int which, const void** data, int64_t* count, int* dtype) {
  if (!m || !data || !count || !dtype) return fail(FS_ERR_ARG, "null argument");
  const fs::Mesh& g = m->m;
  auto I = [&](const std::vector<int>& v) { *data = v.data(); *count = (int64_t)v.size(); *dtype = 0; return FS_OK; };
  auto D = [&](const std::vector<double>& v) { *data = v.data(); *count = (int64_t)v.size(); *dtype = 1; return FS_OK; };
  switch (which) {
    case FS_M_X: return D(g.x);
    case FS_M_TRI: return I(g.tri);
    case FS_M_NN_MAP: return I(g.nn_map);
    case FS_M_NCSN: return I(g.ncsn);
    case FS_M_CSN: return I(g.csn);
    case FS_M_NNSN: return I(g.nnsn);
    case FS_M_NSN: return I(g.nsn);
    case FS_M_ESN: return I(g.esn);
    case FS_M_EDGES: return I(g.edges);
    case FS_M_C2E: return I(g.c2e);
    case FS_M_SUB_EDGE: return I(g.sub_edge);
    case FS_M_SUB_ELEM: return I(g.sub_elem);
    case FS_M_SUB_N: return D(g.sub_n);
    case FS_M_SUB_XM: return D(g.sub_xm);
    case FS_M_SUB_XG: return D(g.sub_xg);
    case FS_M_EL_CENT: return D(g.el_cent);
    case FS_M_BEDGE: return I(g.bedge);
    case FS_M_BEDGE_N: return D(g.bedge_n);
    case FS_M_BEDGE_XG: return D(g.bedge_xg);
    case FS_M_CV_AREA: return D(g.cv_area);
    case FS_M_MOMENTS: return D(g.moments);
    case FS_M_EDGE_WEIGHT: return D(g.edge_weight);
    case FS_M_IA: return I(g.ia);
    case FS_M_IAU: return I(g.iau);
    case FS_M_JA: return I(g.ja);
    case FS_M_COLOR: return I(g.color);
    case FS_M_COLOR_OFF: return I(g.color_off);
    case FS_M_COLOR_ROWS: return I(g.color_rows);
    case FS_M_ST_OFF: return I(g.st_off);
    case FS_M_ST_NODES: return I(g.st_nodes);
    case FS_M_ST_W: return D(g.st_W);
    case FS_M_BBC: return I(g.bbc);
    case FS_M_BSEG_OFF: return I(g.bseg_off);
    case FS_M_BSEG: return I(g.bseg);
    case FS_M_BEDGE_GN: return D(g.bedge_gn);
    case FS_M_QUAD: return I(g.quad);
    default: return fail(FS_ERR_ARG, "unknown mesh array id");
  }
}

void fs_mesh_destroy(fs_mesh* m) { delete m; }

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
int fs_solver_create(const fs_mesh* mh, const fs_params* p, int precision, fs_solver** out) {
  if (!mh || !p || !out) return fail(FS_ERR_ARG, "null argument");
  if (!mh->prepared) return fail(FS_ERR_ARG, "mesh not prepared");
  if (p->order == 3 && mh->order_prepared != 3) return fail(FS_ERR_ARG, "order-3 solver needs a mesh prepared with order 3");
  auto s = std::make_unique<fs_solver>();
  int rc = s->eng.create({&mh->m}, {*p}, precision);
  if (rc) return rc;
  *out = s.release();
  return FS_OK;
}

void fs_solver_destroy(fs_solver* s) { delete s; }

int fs_solver_init_freestream(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  return s->eng.init_freestream(0);
}

int fs_solver_init_restart(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  const fs::Mesh& g = *s->eng.inst[0].mesh;
  if (!g.has_restart) return fail(FS_ERR_ARG, "the solver's mesh carries no restart payload");
  std::vector<double> Q(4 * ((size_t)g.nn + g.ng));
  std::copy(g.restart_Q.begin(), g.restart_Q.end(), Q.begin());
  if (g.restart_ghost_Q.size() == 4 * (size_t)g.ng) {
    std::copy(g.restart_ghost_Q.begin(), g.restart_ghost_Q.end(), Q.begin() + 4 * (size_t)g.nn);
  } else {
    for (int b = 0; b < g.nbe; ++b) {
      int real = g.bedge[5 * (size_t)b], gh = g.nn + g.bedge[5 * (size_t)b + 1];
      for (int v = 0; v < 4; ++v) Q[4 * (size_t)gh + v] = Q[4 * (size_t)real + v];
    }
  }
  int rc = s->eng.set_field(0, FS_F_Q, Q.data(), (int64_t)Q.size());
  if (!rc) rc = s->eng.set_field(0, FS_F_NQ, Q.data(), (int64_t)Q.size());
  return rc;
}

int fs_solver_set(fs_solver* s, int field, const double* host, int64_t n) {
  if (!s || !host) return fail(FS_ERR_ARG, "null argument");
  return s->eng.set_field(0, field, host, n);
}

int fs_solver_get(fs_solver* s, int field, double* host, int64_t n) {
  if (!s || !host) return fail(FS_ERR_ARG, "null argument");
  return s->eng.get_field(0, field, host, n);
}

int fs_solver_use_graph(fs_solver* s, int on) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  s->graph = on != 0;
  return FS_OK;
}

int fs_solver_autotune(fs_solver* s, double* seconds) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  return s->eng.autotune(seconds);
}

int fs_solver_get_tuning(const fs_solver* s, int* values, int* from_cache) {
  if (!s || !values) return fail(FS_ERR_ARG, "null argument");
  s->eng.tuning_values(values);
  if (from_cache) *from_cache = s->eng.tuned_from_cache() ? 1 : 0;
  return FS_OK;
}

int fs_solver_set_tuning(fs_solver* s, const int* values) {
  if (!s || !values) return fail(FS_ERR_ARG, "null argument");
  return s->eng.set_tuning(values);
}

int fs_build_matrix(fs_solver* s, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (int rc = need_f64(s)) return rc;
  if (int rc = s->eng.upload_params_custom(mirror_params(s, ts))) return rc;
  fsk::launch_zero_offdiag(s->eng.dm, s->eng.ds, s->eng.stream);
  fsk::launch_node(s->eng.dm, s->eng.ds, fsk::NODE_EVCON, s->eng.stream);
  return finish(s, true);
}

int fs_eigenvalue_contribution(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (int rc = s->eng.upload_params_custom(mirror_params(s, s->eng.inst[0].ts))) return rc;
  fsk::launch_node(s->eng.dm, s->eng.ds, fsk::NODE_EVCON | fsk::NODE_EVONLY, s->eng.stream);
  return finish(s, true);
}

int fs_set_boundary_values(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  // the boundary pass with the Jacobian off: ghost states (and the
  // boundary fluxes, which build_boundary_residuals reads next)
  fsk::InstParams p = mirror_params(s, s->eng.inst[0].ts);
  p.jac = 0;
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_bedge(s->eng.dm, s->eng.ds, s->eng.stream);
  return finish(s, true);
}

int fs_gradient(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::InstParams p = mirror_params(s, s->eng.inst[0].ts);
  p.grad = 1;
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_gradient(s->eng.dm, s->eng.ds, s->eng.stream);
  return finish(s, false);
}

int fs_derivatives(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (s->eng.inst[0].nn > 0 && s->eng.inst[0].mesh->st_off.empty()) return fail(FS_ERR_ARG, "mesh has no k=2 stencils");
  fsk::InstParams p = mirror_params(s, s->eng.inst[0].ts);
  p.k2 = 1;
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_derivatives(s->eng.dm, s->eng.ds, s->eng.stream);
  return finish(s, false);
}

int fs_limiter(fs_solver* s, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::InstParams p = mirror_params(s, ts);
  // the mirror of compute_limiter_barth / compute_limiter_venkatakrishnan
  // (limit 2) at the points of the current phase
  p.lim = fsk::LIM_ON | (fse::phase_for(s->eng.inst[0].prm, ts).k2 ? fsk::LIM_GAUSS : 0) |
          (s->eng.inst[0].prm.limit == 2 ? fsk::LIM_VENKAT : 0);
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_limiter(s->eng.dm, s->eng.ds, s->eng.stream);
  return finish(s, false);
}

int fs_interior_residual(fs_solver* s, int jupdate, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (int rc = need_f64(s)) return rc;
  fsk::InstParams p = mirror_params(s, ts);
  p.jac = jupdate ? 1 : 0;
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_edge_flux(s->eng.dm, s->eng.ds, s->eng.stream);
  fsk::launch_node(s->eng.dm, s->eng.ds, fsk::NODE_INTERIOR | fsk::NODE_JAC, s->eng.stream);
  return finish(s, false);
}

int fs_boundary_residual(fs_solver* s, int jupdate) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (int rc = need_f64(s)) return rc;
  fsk::InstParams p = mirror_params(s, s->eng.inst[0].ts);
  p.jac = jupdate ? 1 : 0;
  if (int rc = s->eng.upload_params_custom(p)) return rc;
  fsk::launch_bedge(s->eng.dm, s->eng.ds, s->eng.stream);
  fsk::launch_node(s->eng.dm, s->eng.ds, fsk::NODE_BOUNDARY | fsk::NODE_JAC, s->eng.stream);
  return finish(s, true);
}

int fs_linear_solve(fs_solver* s, double* rms) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (int rc = need_f64(s)) return rc;
  fse::Engine& E = s->eng;
  if (int rc = E.upload_params_custom(mirror_params(s, E.inst[0].ts))) return rc;
  fsk::launch_diag_lu(E.dm, E.ds, E.stream);
  fsk::launch_gs_sweeps(E.dm, E.ds, fsk::ColorRanges{E.color_off.data(), E.color_nback.data()}, E.subits, E.stream);
  if (int rc = finish(s, false)) return rc;
  if (rms) {
    // sqrt(sum |RHS - A dQ|^2 / nn) as linear_algebra.C:348-385 reports
    const fs::Mesh& g = *E.inst[0].mesh;
    std::vector<double> L((size_t)E.inst[0].nnz * 16), b((size_t)g.nn * 4), x((size_t)g.nn * 4);
    int rc = E.get_field(0, FS_F_LHS, L.data(), (int64_t)L.size());
    if (!rc) rc = E.get_field(0, FS_F_RHS, b.data(), (int64_t)b.size());
    if (!rc) rc = E.get_field(0, FS_F_DQ, x.data(), (int64_t)x.size());
    if (rc) return rc;
    double acc = 0;
    for (int j = 0; j < g.nn; ++j) {
      double r[4] = {b[4 * j], b[4 * j + 1], b[4 * j + 2], b[4 * j + 3]};
      for (int k = g.ia[j]; k < g.ia[j + 1]; ++k) {
        const double* A = &L[16 * (size_t)k];
        const double* q = &x[4 * (size_t)g.ja[k]];
        for (int v = 0; v < 4; ++v) r[v] -= (A[v * 4] * q[0] + A[v * 4 + 1] * q[1] + A[v * 4 + 2] * q[2] + A[v * 4 + 3] * q[3]);
      }
      acc += r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
    }
    *rms = std::sqrt(acc / (1. * g.nn));
  }
  return FS_OK;
}

int fs_solve(fs_solver* s, int ts, double* dq_norm, double* r2) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fse::Instance& I = s->eng.inst[0];
  I.ts = ts;
  I.active = true;
  int rc = s->eng.step(false, s->graph, s->timing);
  if (rc) return rc;
  if (I.err) return fail(FS_ERR_STATE, fse::device_error_message(I.err));
  if (dq_norm) *dq_norm = std::sqrt(I.dq2);
  if (r2) *r2 = I.r2;
  return FS_OK;
}

int fs_roe_flux_batch(int n, double gamma, const double* d_normals, const double* d_qL, const double* d_qR,
                      double* d_flux, void* stream) {
  if (n < 0 || (n > 0 && (!d_normals || !d_qL || !d_qR || !d_flux))) return fail(FS_ERR_ARG, "bad arguments");
  fsk::launch_roe_batch(n, 0, gamma, d_normals, d_qL, d_qR, d_flux, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FS_OK;
}

int fs_roe_flux_batch_form(int n, int form, double gamma, const double* d_normals, const double* d_qL,
                           const double* d_qR, double* d_flux, void* stream) {
  if (n < 0 || (form != 0 && form != 1) || (n > 0 && (!d_normals || !d_qL || !d_qR || !d_flux)))
    return fail(FS_ERR_ARG, "bad arguments");
  fsk::launch_roe_batch(n, form, gamma, d_normals, d_qL, d_qR, d_flux, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FS_OK;
}

int fs_body_forces(fs_solver* s, int citer, double* forces) {
  if (!s || !forces) return fail(FS_ERR_ARG, "null argument");
  return s->eng.body_forces(&citer, forces);
}

namespace {
// write_surface_cp's nodal branch (grid_io.C:1270-1330): the unique nodes of
// boundary b's segments in order of first appearance, Cp from Q.
int surface_cp_values(fs_solver* s, int b, std::vector<double>* xs, std::vector<double>* cps) {
  const fse::Instance& I = s->eng.inst[0];
  const fs::Mesh& g = *I.mesh;
  if (b < 0 || b >= g.nb) return fail(FS_ERR_ARG, "boundary out of range");
  xs->clear();
  cps->clear();
  const int bc = g.bbc[b];
  if (bc == 0 || (bc % 2 == 0)) return FS_OK;
  std::vector<double> Q(4 * (size_t)(g.nn + g.ng));
  int rc = s->eng.get_field(0, FS_F_Q, Q.data(), (int64_t)Q.size());
  if (rc) return rc;
  const fs_params& p = I.prm;
  const double pinf = 1. / p.gamma, minf = p.mach;
  std::vector<int> seen;
  for (int sg = g.bseg_off[b]; sg < g.bseg_off[b + 1]; ++sg)
    for (int e = 0; e < 2; ++e) {
      const int n = g.bseg[2 * (size_t)sg + e];
      if (std::find(seen.begin(), seen.end(), n) != seen.end()) continue;
      seen.push_back(n);
      const double* q = &Q[4 * (size_t)n];
      const double r = q[0], u = q[1] / r, v = q[2] / r, en = q[3];
      const double P = (p.gamma - 1.) * (en - 0.5 * r * (u * u + v * v));
      xs->push_back(g.x[2 * (size_t)n]);
      cps->push_back((P / pinf - 1.) * (2. / (p.gamma * minf * minf)));
    }
  return FS_OK;
}
}  // namespace

int fs_surface_cp(fs_solver* s, int b, double* x, double* cp, int64_t cap, int64_t* count) {
  if (!s || !count) return fail(FS_ERR_ARG, "null argument");
  std::vector<double> xs, cps;
  int rc = surface_cp_values(s, b, &xs, &cps);
  if (rc) return rc;
  *count = (int64_t)xs.size();
  if (!x || !cp) return FS_OK;
  if (cap < (int64_t)xs.size()) return fail(FS_ERR_ARG, "output arrays too small");
  std::copy(xs.begin(), xs.end(), x);
  std::copy(cps.begin(), cps.end(), cp);
  return FS_OK;
}

int fs_write_surface_cp(fs_solver* s, const char* dir) {
  if (!s || !dir) return fail(FS_ERR_ARG, "null argument");
  const fs::Mesh& g = *s->eng.inst[0].mesh;
  for (int b = 0; b < g.nb; ++b) {
    const int bc = g.bbc[b];
    if (bc == 0 || (bc % 2 == 0)) continue;
    std::vector<double> xs, cps;
    int rc = surface_cp_values(s, b, &xs, &cps);
    if (rc) return rc;
    // the reference prints "%f  %f" of the floats and sorts the file with
    // `sort -g`: numeric order of the first field, then the whole line
    std::vector<std::string> lines;
    char buf[128];
    for (size_t k = 0; k < xs.size(); ++k) {
      std::snprintf(buf, sizeof buf, "%f  %f\n", (double)(float)xs[k], (double)(float)cps[k]);
      lines.push_back(buf);
    }
    std::stable_sort(lines.begin(), lines.end(), [](const std::string& a, const std::string& c) {
      double va = std::strtod(a.c_str(), nullptr), vc = std::strtod(c.c_str(), nullptr);
      if (va != vc) return va < vc;
      return a < c;
    });
    std::string path = std::string(dir) + "/surface_" + std::to_string(b + 1) + "_pressure.dat";
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(FS_ERR_IO, "cannot open " + path);
    for (const auto& l : lines) std::fputs(l.c_str(), fp);
    std::fclose(fp);
  }
  return FS_OK;
}

int fs_solve_steady(fs_solver* s, int max_steps, double tol, int* steps, double* drop) {
  if (!s || max_steps < 1) return fail(FS_ERR_ARG, "bad arguments");
  const fs_params& p = s->eng.inst[0].prm;
  int ts_ref = (p.order == 3) ? p.foits + p.soits + 1 : p.foits + 1;
  double r0 = -1.0, d = 1.0;
  int ts = 1;
  for (; ts <= max_steps; ++ts) {
    double dq, r2;
    int rc = fs_solve(s, ts, &dq, &r2);
    if (rc) return rc;
    double rn = std::sqrt(r2);
    if (ts == ts_ref || (ts > ts_ref && r0 < 0)) r0 = rn;
    if (r0 > 0) {
      d = rn / r0;
      if (d <= tol) break;
    }
  }
  if (steps) *steps = ts > max_steps ? max_steps : ts;
  if (drop) *drop = d;
  return FS_OK;
}

int fs_solver_time_steps(fs_solver* s, int ts_first, int nsteps, fs_timing* t) {
  if (!s || !t || nsteps < 1) return fail(FS_ERR_ARG, "bad arguments");
  fse::Engine& E = s->eng;
  // kernel-group breakdown from events on an eager replay of the same steps
  // is taken only when no graph is requested
  s->timing = !s->graph;
  E.t_edge = E.t_node = E.t_gs = E.t_lu = 0;
  HIP_TRY(hipStreamSynchronize(E.stream));
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < nsteps; ++k) {
    int rc = fs_solve(s, ts_first + k, nullptr, nullptr);
    if (rc) { s->timing = false; return rc; }
  }
  HIP_TRY(hipStreamSynchronize(E.stream));
  auto t1 = std::chrono::steady_clock::now();
  bool timed = s->timing;
  s->timing = false;
  t->step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps;
  t->node_ms = timed ? E.t_node / nsteps : 0;
  t->gs_ms = timed ? E.t_gs / nsteps : 0;
  t->lu_ms = timed ? E.t_lu / nsteps : 0;
  t->edge_ms = timed ? E.t_edge / nsteps : 0;
  t->other_ms = timed ? t->step_ms - t->edge_ms - t->node_ms - t->gs_ms - t->lu_ms : 0;
  t->flux_evals = (int64_t)E.dm.nsub + E.dm.nbe;
  return FS_OK;
}

}  // extern "C"
