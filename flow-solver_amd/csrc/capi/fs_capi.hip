// C-ABI implementation: mesh handles (host setup) and the single-GPU solver
// (device-resident state, one HIP stream per solver).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/fs_setup.hpp"
#include "../kernels/fs_kernels.hpp"
#include "fs_capi.h"

struct fs_mesh {
  fs::Mesh m;
  bool prepared = false;
  int order_prepared = -1;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(FS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

fs::Params to_params(const fs_params* p) {
  fs::Params q;
  q.nsteps = p->nsteps; q.iter = p->iter; q.jufreq = p->jufreq; q.CFL = p->CFL; q.ramp = p->ramp;
  q.gamma = p->gamma; q.alpha = p->alpha; q.mach = p->mach; q.subits = p->subits; q.unsteady = p->unsteady;
  q.tacc = p->tacc; q.mintime = p->mintime; q.method = p->method; q.order = p->order; q.foits = p->foits;
  q.soits = p->soits; q.toits = p->toits; q.isrestart = p->isrestart; q.limit = p->limit; q.reorder = p->reorder;
  q.ic = p->ic; q.p_or_cv = p->p_or_cv; q.start_post_pro = p->start_post_pro; q.post_pro_freq = p->post_pro_freq;
  return q;
}
void from_params(const fs::Params& q, fs_params* p) {
  p->nsteps = q.nsteps; p->iter = q.iter; p->jufreq = q.jufreq; p->CFL = q.CFL; p->ramp = q.ramp;
  p->gamma = q.gamma; p->alpha = q.alpha; p->mach = q.mach; p->subits = q.subits; p->unsteady = q.unsteady;
  p->tacc = q.tacc; p->mintime = q.mintime; p->method = q.method; p->order = q.order; p->foits = q.foits;
  p->soits = q.soits; p->toits = q.toits; p->isrestart = q.isrestart; p->limit = q.limit; p->reorder = q.reorder;
  p->ic = q.ic; p->p_or_cv = q.p_or_cv; p->start_post_pro = q.start_post_pro; p->post_pro_freq = q.post_pro_freq;
}

// device buffer owner
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DBuf() { if (p) (void)hipFree(p); }
};
}  // namespace

struct fs_solver {
  int device = 0;
  int precision = FS_PREC_F64;
  fs_params prm{};
  const fs::Mesh* host = nullptr;
  std::vector<int> color_off;
  std::vector<int> slot_dev;   // reference block slot -> device (color-major) slot
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<DBuf>> bufs;
  fsk::DevMesh dm{};
  fsk::DevState ds{};
  double* d_norms = nullptr;   // 2 doubles
  double* h_norms = nullptr;   // pinned
  int* h_err = nullptr;        // pinned
  double uinf = 0, vinf = 0, pinf = 0;
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  double t_edge = 0, t_node = 0, t_gs = 0, t_lu = 0;

  template <typename T>
  int alloc(T** out, size_t count) {
    auto b = std::make_unique<DBuf>();
    b->bytes = count * sizeof(T);
    if (b->bytes == 0) b->bytes = sizeof(T);
    hipError_t e = hipMalloc(&b->p, b->bytes);
    if (e != hipSuccess) return fail(FS_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    *out = static_cast<T*>(b->p);
    bufs.push_back(std::move(b));
    return FS_OK;
  }
  template <typename T>
  int upload(const T** out, const std::vector<T>& v) {
    T* d = nullptr;
    int rc = alloc(&d, v.size());
    if (rc) return rc;
    if (!v.empty()) HIP_TRY(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = d;
    return FS_OK;
  }
};

namespace {

double cfl_for(const fs_params& p, int citer) {
  // init_memory.C:200-216
  if (p.method == 1) {
    if (citer < p.ramp) return 1. + (1. * (citer - 1)) / (1. * p.ramp) * (p.CFL - 1.);
    return p.CFL;
  }
  if (p.method == 0) return 1.0;
  return p.CFL;
}

struct Phase {
  int recon_flux;  // 0 nodal, 1 linear
  bool gradient;
  bool k2;
  bool limiter;
  int lim_recon;
};

Phase phase_for(const fs_params& p, int ts) {
  Phase ph{0, false, false, false, 1};
  if (p.order == 1) {
    if (ts > p.foits) { ph.recon_flux = 1; ph.gradient = true; ph.limiter = p.limit > 0; }
  } else if (p.order == 3) {
    if (ts > p.foits && ts <= p.foits + p.soits) { ph.recon_flux = 1; ph.gradient = true; }
    else if (ts > p.foits + p.soits) { ph.k2 = true; ph.limiter = p.limit > 0; ph.lim_recon = 3; }
  }
  return ph;
}

fsk::PhysParams phys(const fs_solver* s, int ts) {
  fsk::PhysParams p;
  p.gamma = s->prm.gamma;
  p.uinf = s->uinf;
  p.vinf = s->vinf;
  p.pinf = s->pinf;
  p.cfl = cfl_for(s->prm, ts);
  return p;
}

__global__ void k_fill(double* p, size_t n, double v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill(double* p, size_t n, double v, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

int check_err(fs_solver* s) {
  HIP_TRY(hipMemcpyAsync(s->h_err, s->ds.err, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (*s->h_err) {
    int e = *s->h_err;
    *s->h_err = 0;
    HIP_TRY(hipMemsetAsync(s->ds.err, 0, sizeof(int), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (e & 1) return fail(FS_ERR_STATE, "boundary state outside the characteristic cases (cvbc.C:330-382 exit)");
    return fail(FS_ERR_STATE, "non-positive or non-finite time step (init_memory.C:257-268 exit)");
  }
  return FS_OK;
}

bool supported(const fs_params& p, std::string* why) {
  if (p.order != 1 && p.order != 3) { *why = "order must be 1 (k=1) or 3 (k=2)"; return false; }
  if (p.limit != 0 && p.limit != 1) { *why = "limit must be 0 or 1 (Barth)"; return false; }
  if (p.method != 1 && p.method != 2) { *why = "method must be implicit (1 ramped CFL or 2 fixed CFL)"; return false; }
  if (p.unsteady != 0 || p.tacc != 0) { *why = "unsteady time accuracy is not on this path"; return false; }
  if (p.iter < 1 || p.jufreq < 1 || p.subits < 0) { *why = "iter/jufreq must be >= 1, subits >= 0"; return false; }
  return true;
}

}  // namespace

extern "C" {

const char* fs_last_error(void) { return g_err.c_str(); }

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

int fs_mesh_create_airfoil(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed, fs_mesh** out) {
  if (!out || nwall < 8 || nrad < 3 || !(h0 > 0) || !(rmax > 1) || jitter < 0 || jitter >= 0.5)
    return fail(FS_ERR_ARG, "bad airfoil O-mesh arguments");
  auto h = std::make_unique<fs_mesh>();
  fs::make_airfoil_omesh(nwall, nrad, h0, rmax, jitter, seed, &h->m);
  *out = h.release();
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
  if (precision != FS_PREC_F64 && precision != FS_PREC_MIXED) return fail(FS_ERR_ARG, "unknown precision");
  std::string why;
  if (!supported(*p, &why)) return fail(FS_ERR_UNSUPPORTED, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_HIP, "no HIP device available");
  auto s = std::make_unique<fs_solver>();
  s->precision = precision;
  s->prm = *p;
  const fs::Mesh& g = mh->m;
  s->host = &g;
  HIP_TRY(hipGetDevice(&s->device));
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  {
    double angle = p->alpha * M_PI / 180.;
    s->uinf = p->mach * cos(angle);
    s->vinf = p->mach * sin(angle);
    s->pinf = 1. / p->gamma;
  }
  fsk::DevMesh& d = s->dm;
  d.nn = g.nn; d.ng = g.ng; d.ne = g.ne; d.nsub = g.nsub; d.nbe = g.nbe; d.nnz = (int)g.ja.size();
  d.ncolor = g.ncolor;
  s->color_off = g.color_off;
  int rc = 0;
  std::vector<double> xr(g.x.begin(), g.x.begin() + 2 * (size_t)g.nn);
  std::vector<int> edge_sub(2 * (size_t)g.ne, -1), edge_slot(2 * (size_t)g.ne, 0);
  for (int e = 0; e < g.ne; ++e) {
    for (int k = g.esub_off[e], t = 0; k < g.esub_off[e + 1] && t < 2; ++k, ++t) edge_sub[2 * (size_t)e + t] = g.esub_list[k];
  }
  // color-major block CSR: rows in color order, each row's blocks in the
  // reference order (nsn order, diagonal last); slot_dev maps a reference slot
  // to its device slot
  s->slot_dev.assign(g.ja.size(), 0);
  std::vector<int> rstart(g.nn), iau_d(g.nn), ja_d(g.ja.size()), slot_t_d(g.ja.size());
  {
    int pos = 0;
    for (int t = 0; t < g.nn; ++t) {
      int j = g.color_rows[t];
      rstart[j] = pos;
      for (int k = g.ia[j]; k < g.ia[j + 1]; ++k) s->slot_dev[k] = pos++;
      iau_d[j] = s->slot_dev[g.iau[j]];
    }
    for (size_t k = 0; k < g.ja.size(); ++k) {
      ja_d[s->slot_dev[k]] = g.ja[k];
      slot_t_d[s->slot_dev[k]] = s->slot_dev[g.slot_t[k]];
    }
    for (int e = 0; e < g.ne; ++e) {
      edge_slot[2 * (size_t)e] = s->slot_dev[g.edge_slot_lr[e]];
      edge_slot[2 * (size_t)e + 1] = s->slot_dev[g.edge_slot_rl[e]];
    }
  }
  // pseudo-inverse entry-major: Wt[(o+j)*5 + r] = W[5*o + r*m + j]
  std::vector<double> st_Wt(g.st_W.size());
  for (int i = 0; i + 1 < (int)g.st_off.size(); ++i) {
    int o = g.st_off[i], mm = g.st_off[i + 1] - o;
    for (int j = 0; j < mm; ++j)
      for (int r = 0; r < 5; ++r) st_Wt[5 * ((size_t)o + j) + r] = g.st_W[5 * (size_t)o + (size_t)r * mm + j];
  }
  // node-incidence arrays: everything a node-centric gather reads, in the
  // node's ascending incidence order
  size_t ninc = g.nsub_list.size(), neinc = g.nedge_list.size();
  std::vector<int> sub_inc(2 * (size_t)g.nsub), inc_other(neinc);
  std::vector<double> inc_n(4 * ninc), inc_fc(2 * ninc), inc_xg(6 * ninc), inc_gw(2 * neinc);
  for (size_t k = 0; k < ninc; ++k) {
    int ent = g.nsub_list[k];
    int is = ent & 0x7fffffff;
    sub_inc[2 * (size_t)is + (ent < 0 ? 1 : 0)] = (int)k;
    inc_n[4 * k] = g.sub_n[3 * (size_t)is];
    inc_n[4 * k + 1] = g.sub_n[3 * (size_t)is + 1];
    inc_n[4 * k + 2] = g.sub_n[3 * (size_t)is + 2];
    inc_n[4 * k + 3] = 0.0;
    int el = g.sub_elem[is];
    // face point of Reconstruct (reconstruct.C:181-186): (centroid + midpoint) * 0.5
    inc_fc[2 * k] = (g.el_cent[2 * (size_t)el] + g.sub_xm[2 * (size_t)is]) * 0.5;
    inc_fc[2 * k + 1] = (g.el_cent[2 * (size_t)el + 1] + g.sub_xm[2 * (size_t)is + 1]) * 0.5;
    for (int c = 0; c < 6; ++c) inc_xg[6 * k + c] = g.sub_xg[6 * (size_t)is + c];
  }
  for (size_t k = 0; k < neinc; ++k) {
    int ent = g.nedge_list[k];
    int e = ent & 0x7fffffff;
    bool isR = ent < 0;
    int other = isR ? g.edges[2 * (size_t)e] : g.edges[2 * (size_t)e + 1];
    inc_other[k] = other | (isR ? (int)0x80000000u : 0);
    inc_gw[2 * k] = g.edge_weight[4 * (size_t)e + (isR ? 2 : 0)];
    inc_gw[2 * k + 1] = g.edge_weight[4 * (size_t)e + (isR ? 3 : 1)];
  }
  std::vector<int> breal(g.nbe), bbc(g.nbe);
  for (int i = 0; i < g.nbe; ++i) {
    breal[i] = g.bedge[5 * (size_t)i];
    bbc[i] = g.bbc[g.bedge[5 * (size_t)i + 3]];
  }
  if ((rc = s->upload(&d.x, xr)) || (rc = s->upload(&d.cv_area, g.cv_area)) || (rc = s->upload(&d.edges, g.edges)) ||
      (rc = s->upload(&d.edge_sub, edge_sub)) || (rc = s->upload(&d.edge_slot, edge_slot)) ||
      (rc = s->upload(&d.edge_w, g.edge_weight)) || (rc = s->upload(&d.sub_n, g.sub_n)) ||
      (rc = s->upload(&d.sub_inc, sub_inc)) || (rc = s->upload(&d.inc_n, inc_n)) ||
      (rc = s->upload(&d.inc_fc, inc_fc)) || (rc = s->upload(&d.inc_xg, inc_xg)) ||
      (rc = s->upload(&d.inc_other, inc_other)) || (rc = s->upload(&d.inc_gw, inc_gw)) ||
      (rc = s->upload(&d.nsub_off, g.nsub_off)) || (rc = s->upload(&d.nedge_off, g.nedge_off)) ||
      (rc = s->upload(&d.nbe_off, g.nbe_off)) || (rc = s->upload(&d.nbe_list, g.nbe_list)) ||
      (rc = s->upload(&d.bedge_real, breal)) || (rc = s->upload(&d.bedge_bc, bbc)) ||
      (rc = s->upload(&d.bedge_n, g.bedge_n)) || (rc = s->upload(&d.rstart, rstart)) || (rc = s->upload(&d.ja, ja_d)) ||
      (rc = s->upload(&d.iau, iau_d)) || (rc = s->upload(&d.slot_t, slot_t_d)) ||
      (rc = s->upload(&d.color_rows, g.color_rows)) || (rc = s->upload(&d.st_off, g.st_off)) ||
      (rc = s->upload(&d.st_nodes, g.st_nodes)) || (rc = s->upload(&d.st_Wt, st_Wt)) ||
      (rc = s->upload(&d.moments, g.moments)))
    return rc;
  fsk::DevState& st = s->ds;
  size_t nn = g.nn, ntot = (size_t)g.nn + g.ng;
  if ((rc = s->alloc(&st.Q, ntot * 4)) || (rc = s->alloc(&st.nQ, ntot * 4)) || (rc = s->alloc(&st.dQ, nn * 4)) ||
      (rc = s->alloc(&st.R, nn * 4)) || (rc = s->alloc(&st.RHS, nn * 4)) || (rc = s->alloc(&st.grad, nn * 8)) ||
      (rc = s->alloc(&st.phi, nn)) || (rc = s->alloc(&st.hess, nn * 20)) || (rc = s->alloc(&st.pv, nn * 4)) ||
      (rc = s->alloc(&st.ev_con, nn)) || (rc = s->alloc(&st.dt, nn)) ||
      (rc = s->alloc(&st.LHS, (size_t)d.nnz * 16)) || (rc = s->alloc(&st.LU, nn * 16)) ||
      (rc = s->alloc(&st.inc_flux, (size_t)g.nsub * 8)) || (rc = s->alloc(&st.bflux, (size_t)g.nbe * 4)) ||
      (rc = s->alloc(&st.bjac, (size_t)g.nbe * 16)) || (rc = s->alloc(&st.partial, 2 * 256)) ||
      (rc = s->alloc(&st.err, 1)) || (rc = s->alloc(&s->d_norms, 4)))
    return rc;
  if (precision == FS_PREC_MIXED) {
    if ((rc = s->alloc(&st.LHS32, (size_t)d.nnz * 16)) || (rc = s->alloc(&st.LU32, nn * 16)) ||
        (rc = s->alloc(&st.RHS32, nn * 4)) || (rc = s->alloc(&st.dQ32, nn * 4)))
      return rc;
  }
  HIP_TRY(hipMemset(st.err, 0, sizeof(int)));
  HIP_TRY(hipMemset(st.grad, 0, nn * 8 * sizeof(double)));
  HIP_TRY(hipMemset(st.hess, 0, nn * 20 * sizeof(double)));
  HIP_TRY(hipMemset(st.pv, 0, nn * 4 * sizeof(double)));
  HIP_TRY(hipMemset(st.LHS, 0, (size_t)d.nnz * 16 * sizeof(double)));
  HIP_TRY(hipMemset(st.dQ, 0, nn * 4 * sizeof(double)));
  HIP_TRY(hipMemset(st.R, 0, nn * 4 * sizeof(double)));
  HIP_TRY(hipMemset(st.RHS, 0, nn * 4 * sizeof(double)));
  HIP_TRY(hipHostMalloc((void**)&s->h_norms, 4 * sizeof(double), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&s->h_err, sizeof(int), hipHostMallocDefault));
  *s->h_err = 0;
  fill(st.phi, nn, 1.0, s->stream);
  for (auto& e : s->ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipStreamSynchronize(s->stream));
  *out = s.release();
  return FS_OK;
}

void fs_solver_destroy(fs_solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& e : s->ev) if (e) (void)hipEventDestroy(e);
  if (s->h_norms) (void)hipHostFree(s->h_norms);
  if (s->h_err) (void)hipHostFree(s->h_err);
  s->bufs.clear();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int fs_solver_init_freestream(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  const fs::Mesh& g = *s->host;
  // ic.C:808-840 ic_0 and the ghost copy of ic.C:77-86
  double angle = s->prm.alpha * M_PI / 180.;
  double rho = 1.0, c = 1.0;
  double P = rho * c * c / s->prm.gamma;
  double u = s->prm.mach * cos(angle), v = s->prm.mach * sin(angle);
  double et = P / (s->prm.gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  std::vector<double> Q(4 * ((size_t)g.nn + g.ng));
  for (size_t n = 0; n < (size_t)g.nn + g.ng; ++n) {
    Q[4 * n] = rho; Q[4 * n + 1] = rho * u; Q[4 * n + 2] = rho * v; Q[4 * n + 3] = et;
  }
  HIP_TRY(hipMemcpy(s->ds.Q, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->ds.nQ, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
  return FS_OK;
}

static int field_ptr(fs_solver* s, int field, double** p, int64_t* n) {
  const fs::Mesh& g = *s->host;
  int64_t nn = g.nn, ntot = (int64_t)g.nn + g.ng;
  switch (field) {
    case FS_F_Q: *p = s->ds.Q; *n = ntot * 4; break;
    case FS_F_NQ: *p = s->ds.nQ; *n = ntot * 4; break;
    case FS_F_DQ: *p = s->ds.dQ; *n = nn * 4; break;
    case FS_F_R: *p = s->ds.R; *n = nn * 4; break;
    case FS_F_RHS: *p = s->ds.RHS; *n = nn * 4; break;
    case FS_F_GRAD: *p = s->ds.grad; *n = nn * 8; break;
    case FS_F_PHI: *p = s->ds.phi; *n = nn; break;
    case FS_F_HESS: *p = s->ds.hess; *n = nn * 20; break;
    case FS_F_PV: *p = s->ds.pv; *n = nn * 4; break;
    case FS_F_EVCON: *p = s->ds.ev_con; *n = nn; break;
    case FS_F_DT: *p = s->ds.dt; *n = nn; break;
    case FS_F_LHS: *p = s->ds.LHS; *n = (int64_t)s->dm.nnz * 16; break;
    default: return fail(FS_ERR_ARG, "unknown field id");
  }
  return FS_OK;
}

int fs_solver_set(fs_solver* s, int field, const double* host, int64_t n) {
  if (!s || !host) return fail(FS_ERR_ARG, "null argument");
  double* d; int64_t cnt;
  int rc = field_ptr(s, field, &d, &cnt);
  if (rc) return rc;
  if (n != cnt) return fail(FS_ERR_ARG, "size mismatch: expected " + std::to_string(cnt));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (field == FS_F_LHS) {
    // reference block order -> device color-major order
    std::vector<double> tmp((size_t)n);
    for (size_t k = 0; k < s->slot_dev.size(); ++k)
      std::memcpy(&tmp[16 * (size_t)s->slot_dev[k]], host + 16 * k, 16 * sizeof(double));
    HIP_TRY(hipMemcpy(d, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return FS_OK;
  }
  HIP_TRY(hipMemcpy(d, host, n * sizeof(double), hipMemcpyHostToDevice));
  return FS_OK;
}

int fs_solver_get(fs_solver* s, int field, double* host, int64_t n) {
  if (!s || !host) return fail(FS_ERR_ARG, "null argument");
  if (field == FS_F_LHS32) {
    if (!s->ds.LHS32) return fail(FS_ERR_ARG, "FP32 blocks exist only for a mixed-precision solver");
    int64_t cnt = (int64_t)s->dm.nnz * 16;
    if (n != cnt) return fail(FS_ERR_ARG, "size mismatch: expected " + std::to_string(cnt));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<float> tmp((size_t)n);
    HIP_TRY(hipMemcpy(tmp.data(), s->ds.LHS32, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < s->slot_dev.size(); ++k)
      for (int c = 0; c < 16; ++c) host[16 * k + c] = (double)tmp[16 * (size_t)s->slot_dev[k] + c];
    return FS_OK;
  }
  double* d; int64_t cnt;
  int rc = field_ptr(s, field, &d, &cnt);
  if (rc) return rc;
  if (n != cnt) return fail(FS_ERR_ARG, "size mismatch: expected " + std::to_string(cnt));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (field == FS_F_LHS) {
    std::vector<double> tmp((size_t)n);
    HIP_TRY(hipMemcpy(tmp.data(), d, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < s->slot_dev.size(); ++k)
      std::memcpy(host + 16 * k, &tmp[16 * (size_t)s->slot_dev[k]], 16 * sizeof(double));
    return FS_OK;
  }
  HIP_TRY(hipMemcpy(host, d, n * sizeof(double), hipMemcpyDeviceToHost));
  return FS_OK;
}

int fs_build_matrix(fs_solver* s, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::launch_zero_offdiag(s->dm, s->ds, s->stream);
  fsk::launch_node(s->dm, s->ds, phys(s, ts), fsk::NODE_EVCON, s->stream);
  HIP_TRY(hipGetLastError());
  return check_err(s);
}

int fs_gradient(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::launch_gradient(s->dm, s->ds, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_derivatives(fs_solver* s) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  if (s->dm.nn > 0 && s->host->st_off.empty()) return fail(FS_ERR_ARG, "mesh has no k=2 stencils");
  fsk::launch_derivatives(s->dm, s->ds, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_limiter(fs_solver* s, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  Phase ph = phase_for(s->prm, ts);
  fsk::launch_limiter(s->dm, s->ds, ph.k2 ? 3 : 1, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_interior_residual(fs_solver* s, int jupdate, int ts) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  Phase ph = phase_for(s->prm, ts);
  fsk::PhysParams pp = phys(s, ts);
  fsk::launch_edge_flux(s->dm, s->ds, pp, ph.recon_flux, jupdate, s->stream);
  fsk::launch_node(s->dm, s->ds, pp, fsk::NODE_INTERIOR | (jupdate ? fsk::NODE_JAC : 0), s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_boundary_residual(fs_solver* s, int jupdate) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::PhysParams pp = phys(s, 1);
  fsk::launch_bedge(s->dm, s->ds, pp, jupdate, s->stream);
  fsk::launch_node(s->dm, s->ds, pp, fsk::NODE_BOUNDARY | (jupdate ? fsk::NODE_JAC : 0), s->stream);
  HIP_TRY(hipGetLastError());
  return check_err(s);
}

int fs_linear_solve(fs_solver* s, double* rms) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  fsk::launch_gs(s->dm, s->ds, s->color_off.data(), s->prm.subits, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (rms) {
    // sqrt(sum |RHS - A dQ|^2 / nn) as linear_algebra.C:348-385 reports
    const fs::Mesh& g = *s->host;
    std::vector<double> L((size_t)s->dm.nnz * 16), b((size_t)g.nn * 4), x((size_t)g.nn * 4);
    int rc = fs_solver_get(s, FS_F_LHS, L.data(), (int64_t)L.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(b.data(), s->ds.RHS, b.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(x.data(), s->ds.dQ, x.size() * sizeof(double), hipMemcpyDeviceToHost));
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

// one Newton iteration of the mixed path: FP64 residual (no Jacobian in the
// face-flux pass), FP32 analytic Jacobian and sweeps, FP64 update and norms
static int newton_iteration_mixed(fs_solver* s, int ts, int it, bool record) {
  const fs_params& p = s->prm;
  int jac = ((it - 1) % p.jufreq == 0) ? 1 : 0;
  if (it == 1) jac = 1;
  Phase ph = phase_for(p, ts);
  fsk::PhysParams pp = phys(s, ts);
  hipStream_t st = s->stream;
  if (ph.gradient) fsk::launch_gradient(s->dm, s->ds, st);
  if (ph.k2) fsk::launch_derivatives(s->dm, s->ds, st);
  if (ph.limiter) fsk::launch_limiter(s->dm, s->ds, ph.lim_recon, st);
  else fill(s->ds.phi, s->dm.nn, 1.0, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[0], st));
  fsk::launch_edge_flux(s->dm, s->ds, pp, ph.recon_flux, 0, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[1], st));
  fsk::launch_edge_jac32(s->dm, s->ds, (float)p.gamma, ph.recon_flux, jac, st);
  fsk::launch_bedge(s->dm, s->ds, pp, jac, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[2], st));
  fsk::launch_node_mixed(s->dm, s->ds, pp, jac, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[3], st));
  fsk::launch_diag_lu32(s->dm, s->ds, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[5], st));
  fsk::launch_gs32_sweeps(s->dm, s->ds, s->color_off.data(), p.subits, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[4], st));
  fsk::launch_update_mixed(s->dm, s->ds, st);
  fsk::launch_norms(s->dm, s->ds, s->d_norms, st);
  HIP_TRY(hipGetLastError());
  return FS_OK;
}

// one Newton iteration of solve.C:328-909 on the device
static int newton_iteration(fs_solver* s, int ts, int it, bool record) {
  if (s->precision == FS_PREC_MIXED) return newton_iteration_mixed(s, ts, it, record);
  const fs_params& p = s->prm;
  int jac = ((it - 1) % p.jufreq == 0) ? 1 : 0;
  if (it == 1) jac = 1;
  Phase ph = phase_for(p, ts);
  fsk::PhysParams pp = phys(s, ts);
  hipStream_t st = s->stream;
  if (ph.gradient) fsk::launch_gradient(s->dm, s->ds, st);
  if (ph.k2) fsk::launch_derivatives(s->dm, s->ds, st);
  if (ph.limiter) fsk::launch_limiter(s->dm, s->ds, ph.lim_recon, st);
  else fill(s->ds.phi, s->dm.nn, 1.0, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[0], st));
  fsk::launch_edge_flux(s->dm, s->ds, pp, ph.recon_flux, jac, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[1], st));
  fsk::launch_bedge(s->dm, s->ds, pp, jac, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[2], st));
  fsk::launch_node(s->dm, s->ds, pp,
                   fsk::NODE_EVCON | fsk::NODE_INTERIOR | fsk::NODE_BOUNDARY | fsk::NODE_RHS | (jac ? fsk::NODE_JAC : 0),
                   st);
  if (record) HIP_TRY(hipEventRecord(s->ev[3], st));
  fsk::launch_diag_lu(s->dm, s->ds, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[5], st));
  fsk::launch_gs_sweeps(s->dm, s->ds, s->color_off.data(), p.subits, st);
  if (record) HIP_TRY(hipEventRecord(s->ev[4], st));
  fsk::launch_norms(s->dm, s->ds, s->d_norms, st);
  fsk::launch_update(s->dm, s->ds, st);
  HIP_TRY(hipGetLastError());
  return FS_OK;
}

int fs_solve(fs_solver* s, int ts, double* dq_norm, double* r2) {
  if (!s) return fail(FS_ERR_ARG, "null solver");
  const fs::Mesh& g = *s->host;
  size_t ntot = ((size_t)g.nn + g.ng) * 4;
  hipStream_t st = s->stream;
  // nQ = Q for real and ghost nodes (solve.C:319-325)
  HIP_TRY(hipMemcpyAsync(s->ds.nQ, s->ds.Q, ntot * sizeof(double), hipMemcpyDeviceToDevice, st));
  for (int it = 1; it <= s->prm.iter; ++it) {
    bool rec = s->timing && it == s->prm.iter;
    int rc = newton_iteration(s, ts, it, rec);
    if (rc) return rc;
    if (rec) {
      HIP_TRY(hipEventSynchronize(s->ev[4]));
      float a = 0, b = 0, c = 0, l = 0;
      HIP_TRY(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
      HIP_TRY(hipEventElapsedTime(&b, s->ev[2], s->ev[3]));
      HIP_TRY(hipEventElapsedTime(&l, s->ev[3], s->ev[5]));
      HIP_TRY(hipEventElapsedTime(&c, s->ev[5], s->ev[4]));
      s->t_edge += a; s->t_node += b; s->t_gs += c; s->t_lu += l;
    }
  }
  // Q = nQ for real and ghost nodes (solve.C:913-931)
  HIP_TRY(hipMemcpyAsync(s->ds.Q, s->ds.nQ, ntot * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->h_norms, s->d_norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  int rc = check_err(s);
  if (rc) return rc;
  if (dq_norm) *dq_norm = std::sqrt(s->h_norms[1]);
  if (r2) *r2 = s->h_norms[0];
  return FS_OK;
}

int fs_roe_flux_batch(int n, double gamma, const double* d_normals, const double* d_qL, const double* d_qR,
                      double* d_flux, void* stream) {
  if (n < 0 || (n > 0 && (!d_normals || !d_qL || !d_qR || !d_flux))) return fail(FS_ERR_ARG, "bad arguments");
  fsk::launch_roe_batch(n, gamma, d_normals, d_qL, d_qR, d_flux, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return FS_OK;
}

int fs_solve_steady(fs_solver* s, int max_steps, double tol, int* steps, double* drop) {
  if (!s || max_steps < 1) return fail(FS_ERR_ARG, "bad arguments");
  const fs_params& p = s->prm;
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
  s->timing = true;
  s->t_edge = s->t_node = s->t_gs = s->t_lu = 0;
  HIP_TRY(hipStreamSynchronize(s->stream));
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < nsteps; ++k) {
    int rc = fs_solve(s, ts_first + k, nullptr, nullptr);
    if (rc) { s->timing = false; return rc; }
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  auto t1 = std::chrono::steady_clock::now();
  s->timing = false;
  t->step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps;
  t->node_ms = s->t_node / nsteps;
  t->gs_ms = s->t_gs / nsteps;
  t->lu_ms = s->t_lu / nsteps;
  t->edge_ms = s->t_edge / nsteps;
  t->other_ms = t->step_ms - t->edge_ms - t->node_ms - t->gs_ms - t->lu_ms;
  t->flux_evals = (int64_t)s->dm.nsub + s->dm.nbe;
  return FS_OK;
}

}  // extern "C"
