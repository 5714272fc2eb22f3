// Packing of prepared meshes into one device image and the pseudo-time step
// over every active instance (solve.C:44-931), optionally as one hipGraph.
#include "fs_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

namespace fse {

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

double cfl_for(const fs_params& p, int citer) {
  if (p.method == 1) {
    if (citer < p.ramp) return 1. + (1. * (citer - 1)) / (1. * p.ramp) * (p.CFL - 1.);
    return p.CFL;
  }
  if (p.method == 0) return 1.0;
  return p.CFL;
}

double cfl_for_inst(const Instance& I) {
  if (I.cfl_scale == 1.0 && I.ramp_from == 0) return cfl_for(I.prm, I.ts);
  fs_params q = I.prm;
  q.CFL = std::max(1.0, I.prm.CFL * I.cfl_scale);
  return cfl_for(q, I.ts - I.ramp_from);
}

// order 1: first-order fluxes for foits steps, then limited linear
// reconstruction.  order 3: first order, then soits linear steps, then the
// k=2 phase whose face states are nodal (the reference's residual at k=2 is
// first order; the k=2 derivatives feed only the Gauss-point limiter).
Phase phase_for(const fs_params& p, int ts) {
  Phase ph{0, false, false, 0};
  // solve.C:440-471: the limiter runs in the phase of the final order only;
  // limit 1 is Barth, 2 Venkatakrishnan
  const int kind = p.limit == 2 ? fsk::LIM_VENKAT : 0;
  if (p.order == 1) {
    if (ts > p.foits) {
      ph.recon_flux = 1;
      ph.gradient = true;
      ph.lim = p.limit > 0 ? (fsk::LIM_ON | kind) : 0;
    }
  } else if (p.order == 3) {
    if (ts > p.foits && ts <= p.foits + p.soits) {
      ph.recon_flux = 1;
      ph.gradient = true;
    } else if (ts > p.foits + p.soits) {
      ph.k2 = true;
      ph.lim = p.limit > 0 ? (fsk::LIM_ON | fsk::LIM_GAUSS | kind) : 0;
    }
  }
  return ph;
}

bool supported(const fs_params& p, std::string* why) {
  if (p.order != 1 && p.order != 3) { *why = "order must be 1 (k=1) or 3 (k=2)"; return false; }
  if (p.limit < 0 || p.limit > 2) { *why = "limit must be 0, 1 (Barth) or 2 (Venkatakrishnan)"; return false; }
  if (p.method != 1 && p.method != 2) { *why = "method must be implicit (1 ramped CFL or 2 fixed CFL)"; return false; }
  if (p.unsteady != 0 || p.tacc != 0) { *why = "unsteady time accuracy is not on this path"; return false; }
  if (p.iter < 1 || p.jufreq < 1 || p.subits < 0) { *why = "iter/jufreq must be >= 1, subits >= 0"; return false; }
  if (!(p.gamma > 1.0) || !(p.mach >= 0.0)) { *why = "gamma must exceed 1 and mach be non-negative"; return false; }
  return true;
}

size_t footprint_bytes(const fs::Mesh& g, int prec, bool checkpoints) {
  // block slots: the sliced layout pads each 16-row slice to its largest row
  size_t nn = g.nn, ng = g.ng, ne = g.ne, nsub = g.nsub, nbe = g.nbe, nnz = g.ja.size() + g.nn / 2;
  size_t ninc = g.nsub_list.size(), neinc = g.nedge_list.size(), nst = g.st_nodes.size();
  size_t b = 0;
  // mesh image
  b += nn * (4 * 4 + 8 * 2 + 8 + 8 * 9 + 4 * 4 + 4 * 3) + ne * 4 * 6 + nsub * (8 * 3 + 4 * 2);
  b += ninc * (8 * 12 + 4) + neinc * (4 + 16) + nbe * (4 * 4 + 8 * 15) + nnz * 8 + (nn + 1) * 4 * 4 + nst * (4 + 40);
  // LDS tiles of the k=2 derivatives: slots, neighbour lists, tile offsets
  if (nst > 0) b += nst * (2 + 4) + (nn / fsk::kDerivTileNodes + 2) * 8;
  // state
  b += (nn + ng) * 4 * 8 * 2 + nn * 8 * (4 * 3 + 8 + 1 + 20 + 4 + 2) + nbe * 8 * 22;
  if (prec == FS_PREC_F64) b += ninc * 4 * 8;   // per-incidence face fluxes of the FP64 pass
  if (prec == FS_PREC_MIXED)
    b += nnz * 16 * 4 + nn * (16 * 4 + 4 * 4 * 2) + neinc * (6 * 8 + 4 * 8 + 4);
  else if (prec == FS_PREC_MIXED16)   // FP16 off-diagonal operands, FP32 diagonals, no stored inverse
    b += (nnz - nn) * 16 * 2 + nn * (16 * 4 + 4 * 4 * 2) + neinc * (6 * 8 + 4 * 8 + 4);
  else
    b += nnz * 16 * 8 + nn * 16 * 8;
  b += (nn / fsk::kNormChunkNodes + 2) * (16 + 12);
  if (checkpoints) b += (nn + ng) * 4 * 8 * 2 + 4;
  // the start state of the steps deferring the k=2 fields
  if (nst > 0) b += (nn + ng) * 4 * 8;
  return b;
}

namespace {
bool valid_block(int b) { return b == 64 || b == 128 || b == 256 || b == 512; }
bool valid_tuning(const fsk::Tuning& t) {
  return valid_block(t.gs_block) && (t.gs_lanes == 1 || t.gs_lanes == 4) && valid_block(t.gs32_block) &&
         (t.gs32_pre == 0 || t.gs32_pre == 4 || t.gs32_pre == 6 || t.gs32_pre == 8) && valid_block(t.node_block) &&
         valid_block(t.face_block) && valid_block(t.cell_block) && (t.deriv_lds == 0 || t.deriv_lds == 1) &&
         (t.gs32_scaled == 0 || t.gs32_scaled == 1) && (t.node_fused == 0 || t.node_fused == 1) &&
         (t.gs32_dpp == 0 || t.gs32_dpp == 1);
}
// FS_TUNE_CACHE, else $HOME/.cache/flowsolver/tune_cache.txt
std::string tune_cache_path() {
  if (const char* e = std::getenv("FS_TUNE_CACHE")) return e;
  const char* home = std::getenv("HOME");
  if (!home) return "";
  return std::string(home) + "/.cache/flowsolver/tune_cache.txt";
}

__global__ void k_fill(double* p, size_t n, double v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill(double* p, size_t n, double v, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}
}  // namespace

template <typename T>
int Engine::alloc(T** out, size_t count) {
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return set_error(FS_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  bufs_.push_back({p});
  bytes_ += bytes;
  *out = static_cast<T*>(p);
  return FS_OK;
}

template <typename T>
int Engine::upload(const T** out, const std::vector<T>& v) {
  T* d = nullptr;
  int rc = alloc(&d, v.size());
  if (rc) return rc;
  if (!v.empty()) FSE_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = d;
  return FS_OK;
}

Engine::~Engine() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (h_ip_) (void)hipHostFree(h_ip_);
  if (h_norms_) (void)hipHostFree(h_norms_);
  if (h_err_) (void)hipHostFree(h_err_);
  if (h_fp_) (void)hipHostFree(h_fp_);
  if (h_bforce_) (void)hipHostFree(h_bforce_);
  if (h_rmode_) (void)hipHostFree(h_rmode_);
  for (auto& b : bufs_) (void)hipFree(b.p);
  if (stream) (void)hipStreamDestroy(stream);
}

int Engine::create(const std::vector<const fs::Mesh*>& meshes, const std::vector<fs_params>& prm, int prec) {
  const int B = (int)meshes.size();
  if (B < 1 || (int)prm.size() != B) return set_error(FS_ERR_ARG, "need one parameter set per mesh");
  if (prec != FS_PREC_F64 && prec != FS_PREC_MIXED && prec != FS_PREC_MIXED16)
    return set_error(FS_ERR_ARG, "unknown precision");
  precision = prec;
  subits = prm[0].subits;
  // FP32 sweeps on operands pre-multiplied by the diagonal inverse when a
  // linear solve runs enough sweeps to pay for the pass that forms them
  // (measured: it pays at 40 and 160 sweeps, costs at 20); FP16 operands
  // exist only in that form
  dm.tune.gs32_scaled = (subits >= 30 || prec == FS_PREC_MIXED16) ? 1 : 0;
  iter = prm[0].iter;
  jufreq = prm[0].jufreq;
  for (int i = 0; i < B; ++i) {
    std::string why;
    if (!supported(prm[i], &why)) return set_error(FS_ERR_UNSUPPORTED, "instance " + std::to_string(i) + ": " + why);
    if (prm[i].subits != subits || prm[i].iter != iter || prm[i].jufreq != jufreq)
      return set_error(FS_ERR_ARG, "all instances of a batch share subits, iter and jufreq");
    const fs::Mesh& g = *meshes[i];
    if (g.ng != g.nbe) return set_error(FS_ERR_MESH, "ghost count differs from boundary-edge count");
    if (prm[i].order == 3 && (int)g.st_off.size() != g.nn + 1)
      return set_error(FS_ERR_ARG, "instance " + std::to_string(i) + ": order 3 needs a mesh prepared with order 3");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(FS_ERR_HIP, "no HIP device available");
  FSE_HIP(hipGetDevice(&device));
  FSE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));

  // ---- offsets ----
  inst.resize(B);
  long long NN = 0, NE = 0, NS = 0, NB = 0, NNZ = 0, KINC = 0, EINC = 0, BINC = 0, NST = 0;
  int ncmax = 0;
  for (int i = 0; i < B; ++i) {
    const fs::Mesh& g = *meshes[i];
    Instance& I = inst[i];
    I.mesh = &g;
    I.prm = prm[i];
    double angle = prm[i].alpha * M_PI / 180.;
    I.uinf = prm[i].mach * cos(angle);
    I.vinf = prm[i].mach * sin(angle);
    I.pinf = 1. / prm[i].gamma;
    I.noff = (int)NN; I.nn = g.nn;
    I.eoff = (int)NE; I.ne = g.ne;
    I.soff = (int)NS; I.nsub = g.nsub;
    I.boff = (int)NB; I.nbe = g.nbe;
    I.nnz = (int)g.ja.size();
    I.ncolor = g.ncolor;
    NN += g.nn; NE += g.ne; NS += g.nsub; NB += g.nbe; NNZ += (long long)g.ja.size();
    KINC += (long long)g.nsub_list.size(); EINC += (long long)g.nedge_list.size();
    BINC += (long long)g.nbe_list.size(); NST += (long long)g.st_nodes.size();
    ncmax = std::max(ncmax, g.ncolor);
  }
  if (NN >= (1LL << 30) || NNZ >= (1LL << 31) || KINC >= (1LL << 31) || NST >= (1LL << 31))
    return set_error(FS_ERR_ARG, "batch too large for 32-bit packed indices");
  fsk::DevMesh& d = dm;
  d.nn = (int)NN; d.ng = (int)NB; d.ne = (int)NE; d.nsub = (int)NS; d.nbe = (int)NB; d.nnz = (int)NNZ;
  d.ncolor = ncmax; d.ninst = B;

  // ---- per-node / per-edge / per-subedge / per-bedge arrays ----
  std::vector<int> inst_node(NN), inst_edge(NE), inst_bedge(NB);
  std::vector<double> x(2 * NN), cv(NN), mom(9 * NN);
  std::vector<int> edges(2 * NE), edge_sub(2 * NE, -1), edge_slot(2 * NE);
  std::vector<double> sub_n(3 * NS);
  std::vector<int> sub_edge_d(NS);
  std::vector<int> nsub_off(NN + 1), nedge_off(NN + 1), nbe_off(NN + 1), st_off(NN + 1);
  std::vector<int> sub_inc(2 * NS), inc_other(EINC), inc_sother(KINC), nbe_list(BINC), st_nodes(NST);
  std::vector<double> inc_n(4 * KINC), inc_fc(2 * KINC), inc_xg(6 * KINC), inc_gw(2 * EINC), st_Wt(5 * NST);
  const bool mixed = prec != FS_PREC_F64;
  std::vector<double> einc_geo(mixed ? 6 * EINC : 0), einc_fc(mixed ? 4 * EINC : 0);
  std::vector<int> einc_slot(mixed ? EINC : 0);
  std::vector<int> breal(NB), bbc(NB);
  std::vector<double> bn(3 * NB), bxg(6 * NB), bgn(6 * NB);
  long long kinc = 0, einc = 0, binc = 0, sto = 0;
  for (int i = 0; i < B; ++i) {
    const fs::Mesh& g = *meshes[i];
    const Instance& I = inst[i];
    const int no = I.noff, eo = I.eoff, so = I.soff, bo = I.boff;
    for (int n = 0; n < g.nn; ++n) {
      inst_node[no + n] = i;
      x[2 * (size_t)(no + n)] = g.x[2 * (size_t)n];
      x[2 * (size_t)(no + n) + 1] = g.x[2 * (size_t)n + 1];
      cv[no + n] = g.cv_area[n];
      for (int c = 0; c < 9; ++c) mom[9 * (size_t)(no + n) + c] = g.moments[9 * (size_t)n + c];
      nsub_off[no + n] = (int)(kinc + g.nsub_off[n]);
      nedge_off[no + n] = (int)(einc + g.nedge_off[n]);
      nbe_off[no + n] = (int)(binc + g.nbe_off[n]);
      st_off[no + n] = (int)(sto + (g.st_off.empty() ? 0 : g.st_off[n]));
    }
    for (int e = 0; e < g.ne; ++e) {
      inst_edge[eo + e] = i;
      edges[2 * (size_t)(eo + e)] = g.edges[2 * (size_t)e] + no;
      edges[2 * (size_t)(eo + e) + 1] = g.edges[2 * (size_t)e + 1] + no;
      for (int k = g.esub_off[e], t = 0; k < g.esub_off[e + 1] && t < 2; ++k, ++t)
        edge_sub[2 * (size_t)(eo + e) + t] = g.esub_list[k] + so;
    }
    for (int s = 0; s < g.nsub; ++s) {
      for (int c = 0; c < 3; ++c) sub_n[3 * (size_t)(so + s) + c] = g.sub_n[3 * (size_t)s + c];
      sub_edge_d[so + s] = g.sub_edge[s] + eo;
    }
    // subedge incidences: normals, face points (reconstruct.C:181-186) and
    // Gauss points, in each node's ascending subedge order
    for (size_t k = 0; k < g.nsub_list.size(); ++k) {
      int ent = g.nsub_list[k];
      int is = ent & 0x7fffffff;
      size_t K = (size_t)kinc + k;
      sub_inc[2 * (size_t)(so + is) + (ent < 0 ? 1 : 0)] = (int)K;
      {
        const int e = g.sub_edge[is];
        const int other = (ent < 0 ? g.edges[2 * (size_t)e] : g.edges[2 * (size_t)e + 1]) + no;
        inc_sother[K] = other | (ent < 0 ? (int)0x80000000u : 0);
      }
      for (int c = 0; c < 3; ++c) inc_n[4 * K + c] = g.sub_n[3 * (size_t)is + c];
      inc_n[4 * K + 3] = 0.0;
      int el = g.sub_elem[is];
      inc_fc[2 * K] = (g.el_cent[2 * (size_t)el] + g.sub_xm[2 * (size_t)is]) * 0.5;
      inc_fc[2 * K + 1] = (g.el_cent[2 * (size_t)el + 1] + g.sub_xm[2 * (size_t)is + 1]) * 0.5;
      for (int c = 0; c < 6; ++c) inc_xg[6 * K + c] = g.sub_xg[6 * (size_t)is + c];
    }
    for (size_t k = 0; k < g.nedge_list.size(); ++k) {
      int ent = g.nedge_list[k];
      int e = ent & 0x7fffffff;
      bool isR = ent < 0;
      int other = (isR ? g.edges[2 * (size_t)e] : g.edges[2 * (size_t)e + 1]) + no;
      size_t K = (size_t)einc + k;
      inc_other[K] = other | (isR ? (int)0x80000000u : 0);
      inc_gw[2 * K] = g.edge_weight[4 * (size_t)e + (isR ? 2 : 0)];
      inc_gw[2 * K + 1] = g.edge_weight[4 * (size_t)e + (isR ? 3 : 1)];
      if (mixed) {
        // the edge's subedges: normal and length, face point (reconstruct.C:181-186)
        for (int t = 0, j = g.esub_off[e]; t < 2; ++t, ++j) {
          double* geo = &einc_geo[6 * K + 3 * t];
          double* fc = &einc_fc[4 * K + 2 * t];
          if (j < g.esub_off[e + 1]) {
            const int sb = g.esub_list[j], el = g.sub_elem[sb];
            for (int c = 0; c < 3; ++c) geo[c] = g.sub_n[3 * (size_t)sb + c];
            fc[0] = (g.el_cent[2 * (size_t)el] + g.sub_xm[2 * (size_t)sb]) * 0.5;
            fc[1] = (g.el_cent[2 * (size_t)el + 1] + g.sub_xm[2 * (size_t)sb + 1]) * 0.5;
          } else {
            geo[0] = geo[1] = geo[2] = 0.0;
            fc[0] = fc[1] = 0.0;
          }
        }
      }
    }
    for (size_t k = 0; k < g.nbe_list.size(); ++k) nbe_list[(size_t)binc + k] = g.nbe_list[k] + bo;
    for (int b = 0; b < g.nbe; ++b) {
      inst_bedge[bo + b] = i;
      breal[bo + b] = g.bedge[5 * (size_t)b] + no;
      bbc[bo + b] = g.bbc[g.bedge[5 * (size_t)b + 3]];
      for (int c = 0; c < 3; ++c) bn[3 * (size_t)(bo + b) + c] = g.bedge_n[3 * (size_t)b + c];
      for (int c = 0; c < 6; ++c) bxg[6 * (size_t)(bo + b) + c] = g.bedge_xg[6 * (size_t)b + c];
      for (int c = 0; c < 6; ++c) bgn[6 * (size_t)(bo + b) + c] = g.bedge_gn[6 * (size_t)b + c];
    }
    // pseudo-inverse entry-major: Wt[(o+j)*5 + r] = W[5*o + r*m + j]
    for (int n = 0; n + 1 < (int)g.st_off.size(); ++n) {
      int o = g.st_off[n], mm = g.st_off[n + 1] - o;
      for (int j = 0; j < mm; ++j) {
        st_nodes[(size_t)sto + o + j] = g.st_nodes[o + j] + no;
        for (int r = 0; r < 5; ++r)
          st_Wt[5 * ((size_t)sto + o + j) + r] = g.st_W[5 * (size_t)o + (size_t)r * mm + j];
      }
    }
    kinc += (long long)g.nsub_list.size();
    einc += (long long)g.nedge_list.size();
    binc += (long long)g.nbe_list.size();
    sto += (long long)g.st_nodes.size();
  }
  nsub_off[NN] = (int)kinc;
  nedge_off[NN] = (int)einc;
  d.maxdeg = 0;
  for (long long n = 0; n < NN; ++n) d.maxdeg = std::max(d.maxdeg, nedge_off[n + 1] - nedge_off[n]);
  if (prec == FS_PREC_MIXED16 && !fsk::resid_jac32_fused_fits(d))
    return set_error(FS_ERR_UNSUPPORTED, "FP16 sweep operands: a node has too many edges (" +
                                             std::to_string(d.maxdeg) + ") for the node pass's LDS rows");
  nbe_off[NN] = (int)binc;
  st_off[NN] = (int)sto;

  // ---- LDS tiles of the k=2 derivative pass ----
  // Consecutive runs of kDerivTileNodes nodes (fewer when their distinct
  // stencil neighbours would exceed kDerivTileMaxU); every stencil entry gets
  // the slot of its node in its tile's neighbour list.
  std::vector<int> dt_off, dt_uoff, dt_unodes;
  std::vector<unsigned short> st_loc;
  if (NST > 0) {
    st_loc.resize(NST);
    std::vector<int> stamp(NN, -1), slot(NN, 0);
    int tile = 0, tstart = 0, nu = 0;
    dt_off.push_back(0);
    dt_uoff.push_back(0);
    for (int n = 0; n < NN; ++n) {
      int add = 0;
      for (int j = st_off[n]; j < st_off[n + 1]; ++j) add += stamp[st_nodes[j]] != tile;
      if (n - tstart == fsk::kDerivTileNodes || nu + add > fsk::kDerivTileMaxU) {
        ++tile;
        tstart = n;
        nu = 0;
        dt_off.push_back(n);
        dt_uoff.push_back((int)dt_unodes.size());
      }
      for (int j = st_off[n]; j < st_off[n + 1]; ++j) {
        const int nb = st_nodes[j];
        if (stamp[nb] != tile) {
          stamp[nb] = tile;
          slot[nb] = nu++;
          dt_unodes.push_back(nb);
        }
        st_loc[j] = (unsigned short)slot[nb];
      }
    }
    dt_off.push_back(NN);
    dt_uoff.push_back((int)dt_unodes.size());
  }
  d.ndtile = dt_off.empty() ? 0 : (int)dt_off.size() - 1;

  // ---- rows in color order over all instances ----
  // Global color c holds the color-c rows of every instance, instance by
  // instance; rows of instances whose last color is c come after the others
  // (backward launches of c stop before them).
  std::vector<int> color_rows(NN), pos_inst(NN), pos_of(NN), pdeg(NN);
  color_off.assign(ncmax + 1, 0);
  color_nback.assign(ncmax, 0);
  {
    long long t = 0;
    for (int c = 0; c < ncmax; ++c) {
      color_off[c] = (int)t;
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < B; ++i) {
          const fs::Mesh& g = *meshes[i];
          Instance& I = inst[i];
          if (c >= g.ncolor || ((g.ncolor - 1 == c) != (pass == 1))) continue;
          for (int tt = g.color_off[c]; tt < g.color_off[c + 1]; ++tt) {
            int j = g.color_rows[tt];
            if (g.iau[j] != g.ia[j + 1] - 1) return set_error(FS_ERR_MESH, "diagonal block is not last in its row");
            int J = I.noff + j;
            color_rows[t] = J;
            pos_inst[t] = i;
            pos_of[J] = (int)t;
            pdeg[t] = g.ia[j + 1] - g.ia[j] - 1;
            ++t;
          }
        }
        if (pass == 0) color_nback[c] = (int)(t - color_off[c]);
      }
    }
    color_off[ncmax] = (int)t;
    if (t != NN) return set_error(FS_ERR_MESH, "coloring does not cover every row exactly once");
  }
  // ---- sliced block slots ----
  // slice s = positions 16s..16s+15; its off-diagonal blocks occupy
  // 16 * (largest row degree of the slice) slots starting at sbase[s]
  // (when every slice padded to the largest row degree costs at most a
  // quarter more slots, all slices get that width: a sweep then computes a
  // row's first slot instead of loading it)
  const long long nslice = (NN + 15) / 16;
  std::vector<int> sbase(nslice + 1, 0);
  {
    long long acc = 0;
    int wmax = 0;
    for (long long sl = 0; sl < nslice; ++sl) {
      int mx = 0;
      for (long long t = 16 * sl; t < std::min(NN, 16 * sl + 16); ++t) mx = std::max(mx, pdeg[t]);
      sbase[sl] = (int)acc;
      acc += 16LL * mx;
      wmax = std::max(wmax, mx);
    }
    d.slice_w = 0;
    if (16LL * wmax * nslice <= acc + acc / 4) {
      d.slice_w = wmax;
      acc = 16LL * wmax * nslice;
      for (long long sl = 0; sl <= nslice; ++sl) sbase[sl] = (int)std::min<long long>(16LL * wmax * sl, acc);
    }
    if (acc + NN + 16LL * fsk::kJaPad >= (1LL << 31))
      return set_error(FS_ERR_ARG, "batch too large for 32-bit block slots");
    sbase[nslice] = (int)acc;
  }
  const long long NOFF = sbase[nslice], NSLOT = NOFF + NN;
  std::vector<int> ja_d(NOFF + 16LL * fsk::kJaPad, 0), iau_d(NN), roff(NN + 1, 0), tslot;
  tslot.reserve(NNZ);
  for (auto& I : inst) I.slot_dev.assign(I.nnz, 0);
  for (int i = 0; i < B; ++i) {
    const fs::Mesh& g = *meshes[i];
    Instance& I = inst[i];
    for (int j = 0; j < g.nn; ++j) {
      int J = I.noff + j;
      int t = pos_of[J];
      long long base = (long long)sbase[t >> 4] + (t & 15);
      for (int k = g.ia[j]; k < g.ia[j + 1] - 1; ++k) {
        long long slot = base + 16LL * (k - g.ia[j]);
        I.slot_dev[k] = (int)slot;
        // the mixed sweeps index their vectors by color position
        ja_d[slot] = mixed ? pos_of[g.ja[k] + I.noff] : g.ja[k] + I.noff;
      }
      I.slot_dev[g.iau[j]] = (int)(NOFF + t);
      iau_d[J] = (int)(NOFF + t);
    }
  }
  for (int i = 0; i < B; ++i) {
    const fs::Mesh& g = *meshes[i];
    const Instance& I = inst[i];
    for (int j = 0; j < g.nn; ++j) {
      roff[I.noff + j] = (int)tslot.size();
      for (int k = g.ia[j]; k < g.ia[j + 1] - 1; ++k) tslot.push_back(I.slot_dev[g.slot_t[k]]);
    }
    for (int e = 0; e < g.ne; ++e) {
      edge_slot[2 * (size_t)(I.eoff + e)] = I.slot_dev[g.edge_slot_lr[e]];
      edge_slot[2 * (size_t)(I.eoff + e) + 1] = I.slot_dev[g.edge_slot_rl[e]];
    }
    if (mixed) {
      // block (this node, other node) of each edge incidence: (L,R) seen from
      // L, (R,L) seen from R
      const size_t e0 = (size_t)nedge_off[I.noff];
      for (size_t k = 0; k < g.nedge_list.size(); ++k) {
        const int ent = g.nedge_list[k];
        const int e = ent & 0x7fffffff;
        einc_slot[e0 + k] = edge_slot[2 * (size_t)(I.eoff + e) + (ent < 0 ? 1 : 0)];
      }
    }
  }
  roff[NN] = (int)tslot.size();
  d.nnz = (int)NSLOT;

  // ---- norm chunks ----
  std::vector<int> ch_inst, ch_b, ch_e, inst_ch0(B + 1);
  for (int i = 0; i < B; ++i) {
    inst_ch0[i] = (int)ch_inst.size();
    for (int n0 = 0; n0 < inst[i].nn; n0 += fsk::kNormChunkNodes) {
      ch_inst.push_back(i);
      ch_b.push_back(inst[i].noff + n0);
      ch_e.push_back(inst[i].noff + std::min(inst[i].nn, n0 + fsk::kNormChunkNodes));
    }
  }
  inst_ch0[B] = (int)ch_inst.size();
  nc.nchunk = (int)ch_inst.size();

  {
    // each instance's positions within each color
    color_segs_.assign(ncmax, {});
    for (int c = 0; c < ncmax; ++c)
      for (int t = color_off[c]; t < color_off[c + 1]; ++t) {
        auto& v = color_segs_[c];
        if (!v.empty() && v.back().inst == pos_inst[t] && v.back().t1 == t)
          ++v.back().t1;
        else
          v.push_back({t, t + 1, pos_inst[t]});
      }
    dt_off_h_ = dt_off;
    // work-list capacity: every chunk of every sweep range, of the nodes,
    // edges and positions (one more per instance and color for the
    // boundaries), and every tile
    long long cap = 64 + (long long)dt_off.size();
    for (int c = 0; c < ncmax; ++c)
      cap += (color_off[c + 1] - color_off[c] + fsk::kGsChunk - 1) / fsk::kGsChunk +
             (color_nback[c] + fsk::kGsChunk - 1) / fsk::kGsChunk;
    cap += 2 * (NN / fsk::kWorkChunk) + NE / fsk::kWorkChunk + 3LL * B + (long long)B * ncmax + 8;
    int rc0 = alloc(&d_chunks_, (size_t)cap);
    if (rc0) return rc0;
  }
  int rc = 0;
  if ((rc = upload(&d.inst_node, inst_node)) || (rc = upload(&d.inst_edge, inst_edge)) ||
      (rc = upload(&d.inst_bedge, inst_bedge)) || (rc = upload(&d.pos_inst, pos_inst)) || (rc = upload(&d.x, x)) ||
      (rc = upload(&d.cv_area, cv)) || (rc = upload(&d.edges, edges)) || (rc = upload(&d.edge_sub, edge_sub)) ||
      (rc = upload(&d.edge_slot, edge_slot)) || (rc = upload(&d.sub_n, sub_n)) || (rc = upload(&d.sub_edge, sub_edge_d)) ||
      (rc = upload(&d.nsub_off, nsub_off)) || (rc = upload(&d.nedge_off, nedge_off)) ||
      (rc = upload(&d.sub_inc, sub_inc)) || (rc = upload(&d.inc_n, inc_n)) || (rc = upload(&d.inc_fc, inc_fc)) ||
      (rc = upload(&d.inc_xg, inc_xg)) || (rc = upload(&d.inc_other, inc_other)) ||
      (rc = upload(&d.inc_sother, inc_sother)) ||
      (rc = upload(&d.inc_gw, inc_gw)) || (rc = upload(&d.nbe_off, nbe_off)) ||
      (rc = upload(&d.einc_geo, einc_geo)) || (rc = upload(&d.einc_fc, einc_fc)) ||
      (rc = upload(&d.einc_slot, einc_slot)) ||
      (rc = upload(&d.nbe_list, nbe_list)) || (rc = upload(&d.bedge_real, breal)) ||
      (rc = upload(&d.bedge_bc, bbc)) || (rc = upload(&d.bedge_n, bn)) || (rc = upload(&d.bedge_xg, bxg)) ||
      (rc = upload(&d.bedge_gn, bgn)) || (rc = upload(&d.roff, roff)) ||
      (rc = upload(&d.tslot, tslot)) || (rc = upload(&d.ja, ja_d)) || (rc = upload(&d.iau, iau_d)) ||
      (rc = upload(&d.pdeg, pdeg)) || (rc = upload(&d.sbase, sbase)) ||
      (rc = upload(&d.color_rows, color_rows)) || (rc = upload(&d.st_off, st_off)) ||
      (rc = upload(&d.st_nodes, st_nodes)) || (rc = upload(&d.st_Wt, st_Wt)) || (rc = upload(&d.moments, mom)) ||
      (rc = upload(&d.dt_off, dt_off)) || (rc = upload(&d.dt_uoff, dt_uoff)) || (rc = upload(&d.dt_unodes, dt_unodes)) ||
      (rc = upload(&d.st_loc, st_loc)) ||
      (rc = upload(&nc.chunk_inst, ch_inst)) || (rc = upload(&nc.chunk_begin, ch_b)) ||
      (rc = upload(&nc.chunk_end, ch_e)) || (rc = upload(&nc.inst_chunk0, inst_ch0)))
    return rc;

  // ---- state ----
  fsk::DevState& st = ds;
  size_t nn = NN, ntot = (size_t)NN + NB;
  if ((rc = alloc(&st.Q, ntot * 4)) || (rc = alloc(&st.nQ, ntot * 4)) || (rc = alloc(&st.dQ, nn * 4)) ||
      (rc = alloc(&st.R, nn * 4)) || (rc = alloc(&st.RHS, nn * 4)) || (rc = alloc(&st.grad, nn * 8)) ||
      (rc = alloc(&st.phi, nn)) || (rc = alloc(&st.hess, nn * 20)) || (rc = alloc(&st.pv, nn * 4)) ||
      (rc = alloc(&st.ev_con, nn)) || (rc = alloc(&st.dt, nn)) ||
      (prec == FS_PREC_F64 && (rc = alloc(&st.inc_flux, (size_t)KINC * 4))) ||
      (rc = alloc(&st.bflux, (size_t)NB * 4)) || (rc = alloc(&st.bjac, (size_t)NB * 16)) ||
      (rc = alloc(&st.bforce, (size_t)NB * 2)) || (rc = alloc(&d_fp_, (size_t)B)) ||
      (rc = alloc(&st.err, (size_t)B)) || (rc = alloc(&st.inst_norms, 2 * (size_t)B)) ||
      (rc = alloc(&st.chunk_part, 2 * (size_t)std::max(1, nc.nchunk))) ||
      (rc = alloc(&d_ip_, (size_t)iter * B)))
    return rc;
  if (mixed) {
    if ((rc = alloc(&st.pos_act, nn))) return rc;
    FSE_HIP(hipMemset(st.pos_act, 0, nn));
  }
  k2_stale_.assign(B, 0);
  phi_one_.assign(B, 1);   // phi starts at 1 everywhere (filled below)
  lim_set_.assign(B, 2);   // no list built yet
  if (B > 1 && (rc = alloc(&d_lim_chunks_, nn / fsk::kWorkChunk + (size_t)B + 2))) return rc;
  bool any_k2 = false;
  for (int i = 0; i < B; ++i) any_k2 = any_k2 || prm[i].order == 3;
  if (any_k2 && iter == 1 && (rc = alloc(&qpre_, ntot * 4))) return rc;
  if (precision == FS_PREC_MIXED) {
    if ((rc = alloc(&st.LHS32, (size_t)NSLOT * 16)) || (rc = alloc(&st.LU32, nn * 16)) ||
        (rc = alloc(&st.RHS32, nn * 4)) || (rc = alloc(&st.dQ32, nn * 4)))
      return rc;
    st.D32 = st.LHS32 + 16 * (size_t)NOFF;
  } else if (precision == FS_PREC_MIXED16) {
    if ((rc = alloc(&st.LHS16, (size_t)NOFF * 16)) || (rc = alloc(&st.D32, nn * 16)) ||
        (rc = alloc(&st.RHS32, nn * 4)) || (rc = alloc(&st.dQ32, nn * 4)))
      return rc;
    // padding slots of the sliced layout are read by no sweep; zero them once
    FSE_HIP(hipMemset(st.LHS16, 0, (size_t)NOFF * 16 * sizeof(unsigned short)));
  } else {
    if ((rc = alloc(&st.LHS, (size_t)NSLOT * 16)) || (rc = alloc(&st.LU, nn * 16))) return rc;
    FSE_HIP(hipMemset(st.LHS, 0, (size_t)NSLOT * 16 * sizeof(double)));
  }
  st.ip = d_ip_;
  FSE_HIP(hipMemset(st.err, 0, (size_t)B * sizeof(int)));
  FSE_HIP(hipMemset(st.grad, 0, nn * 8 * sizeof(double)));
  FSE_HIP(hipMemset(st.hess, 0, nn * 20 * sizeof(double)));
  FSE_HIP(hipMemset(st.pv, 0, nn * 4 * sizeof(double)));
  FSE_HIP(hipMemset(st.dQ, 0, nn * 4 * sizeof(double)));
  FSE_HIP(hipMemset(st.R, 0, nn * 4 * sizeof(double)));
  FSE_HIP(hipMemset(st.RHS, 0, nn * 4 * sizeof(double)));
  FSE_HIP(hipMemset(st.inst_norms, 0, 2 * (size_t)B * sizeof(double)));
  FSE_HIP(hipHostMalloc((void**)&h_ip_, (size_t)iter * B * sizeof(fsk::InstParams), hipHostMallocDefault));
  FSE_HIP(hipHostMalloc((void**)&h_norms_, 2 * (size_t)B * sizeof(double), hipHostMallocDefault));
  FSE_HIP(hipHostMalloc((void**)&h_err_, (size_t)B * sizeof(int), hipHostMallocDefault));
  FSE_HIP(hipHostMalloc((void**)&h_fp_, (size_t)B * sizeof(fsk::ForceParams), hipHostMallocDefault));
  FSE_HIP(hipHostMalloc((void**)&h_bforce_, 2 * std::max<size_t>(1, (size_t)NB) * sizeof(double),
                        hipHostMallocDefault));
  std::memset(h_ip_, 0, (size_t)iter * B * sizeof(fsk::InstParams));
  std::memset(h_err_, 0, (size_t)B * sizeof(int));
  fill(st.phi, nn, 1.0, stream);
  for (auto& e : ev_) FSE_HIP(hipEventCreate(&e));
  FSE_HIP(hipStreamSynchronize(stream));
  // launch shapes of an earlier tuning run with the same key
  {
    std::string path = tune_cache_path(), key = tune_key(), line;
    std::ifstream in(path);
    while (in && std::getline(in, line)) {
      std::istringstream ls(line);
      std::string k;
      fsk::Tuning t;
      if (ls >> k >> t.gs_block >> t.gs_lanes >> t.gs32_block >> t.gs32_pre >> t.node_block >> t.face_block >>
          t.cell_block >> t.deriv_lds >> t.node_fused >> t.gs32_dpp && k == key) {
        // the sweep form is the engine's own (it changes FP32 rounding)
        t.gs32_scaled = dm.tune.gs32_scaled;
        if (!valid_tuning(t)) continue;
        dm.tune = t;
        tuned_from_cache_ = true;
      }
    }
  }
  return FS_OK;
}

int Engine::init_freestream(int which) {
  FSE_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < ninst(); ++i) {
    if (which >= 0 && i != which) continue;
    const Instance& I = inst[i];
    double angle = I.prm.alpha * M_PI / 180.;
    double rho = 1.0, c = 1.0;
    double P = rho * c * c / I.prm.gamma;
    double u = I.prm.mach * cos(angle), v = I.prm.mach * sin(angle);
    double et = P / (I.prm.gamma - 1.0) + 0.5 * rho * (u * u + v * v);
    std::vector<double> Q(4 * ((size_t)I.nn + I.nbe));
    for (size_t n = 0; n < (size_t)I.nn + I.nbe; ++n) {
      Q[4 * n] = rho; Q[4 * n + 1] = rho * u; Q[4 * n + 2] = rho * v; Q[4 * n + 3] = et;
    }
    int rc = set_field(i, FS_F_Q, Q.data(), (int64_t)Q.size());
    if (!rc) rc = set_field(i, FS_F_NQ, Q.data(), (int64_t)Q.size());
    if (rc) return rc;
  }
  return FS_OK;
}

fsk::InstParams Engine::params_for(int i, int it) const {
  const Instance& I = inst[i];
  fsk::InstParams p{};
  p.gamma = I.prm.gamma;
  p.uinf = I.uinf;
  p.vinf = I.vinf;
  p.pinf = I.pinf;
  p.cfl = cfl_for_inst(I);
  Phase ph = phase_for(I.prm, I.ts);
  p.recon_flux = ph.recon_flux;
  p.grad = ph.gradient ? 1 : 0;
  p.k2 = ph.k2 ? 1 : 0;
  p.lim = ph.lim;
  if (I.freeze > 0 && (ph.lim & fsk::LIM_ON)) {
    // the limiter runs from step start + 1 (solve.C:440-471)
    const int start = I.prm.order == 3 ? I.prm.foits + I.prm.soits : I.prm.foits;
    if (I.ts - start > I.freeze) p.lim |= fsk::LIM_FROZEN;
  }
  // solve.C:335-338: update on the first iteration and every jufreq-th
  p.jac = (it == 1 || (it - 1) % jufreq == 0) ? 1 : 0;
  p.active = I.active ? 1 : 0;
  if (k2_deferred(I)) {
    p.k2_deferred = 1;
    p.k2 = 0;
    p.lim |= fsk::LIM_FROZEN;
  }
  // a limiter pass without a limiter writes phi = 1: nothing to do where
  // every phi of the instance is 1 already
  if (!(p.lim & (fsk::LIM_ON | fsk::LIM_FROZEN)) && phi_one_[i]) p.lim |= fsk::LIM_FROZEN;
  return p;
}

// what a limiter pass with mode `lim` leaves in instance i's phi
void Engine::note_phi(int i, int lim) {
  if (lim & fsk::LIM_FROZEN) return;
  phi_one_[i] = (lim & fsk::LIM_ON) ? 0 : 1;
}

bool Engine::k2_deferred(const Instance& I) const {
  return defer_k2 && !k2_trial_ && qpre_ && iter == 1 && I.freeze == 0 && phase_for(I.prm, I.ts).k2;
}

// after upload_params of a step: which instances' k=2 fields that step leaves
// to ensure_k2_fields (a step that evaluates them makes them current)
void Engine::mark_k2_stale() {
  for (int i = 0; i < ninst(); ++i)
    if (inst[i].active) {
      k2_stale_[i] = h_ip_[i].k2_deferred ? 1 : 0;
      note_phi(i, h_ip_[i].lim);
    }
}

// after upload_params of a step
int Engine::refresh_limiter_list() {
  if (!d_lim_chunks_) return FS_OK;
  const int B = ninst();
  std::vector<char> need(B, 0);
  long long nodes = 0;
  for (int i = 0; i < B; ++i)
    if (inst[i].active && !(h_ip_[i].lim & fsk::LIM_FROZEN)) {
      need[i] = 1;
      nodes += inst[i].nn;
    }
  if (need == lim_set_) return FS_OK;
  lim_set_ = need;
  if (4 * nodes > 3LL * dm.nn) {
    dm.lim_chunks = nullptr;
    dm.n_lim_chunks = 0;
  } else {
    lim_h_.clear();
    for (int i = 0; i < B; ++i) {
      const Instance& I = inst[i];
      if (!need[i] || I.nn == 0) continue;
      for (int c = I.noff / fsk::kWorkChunk; c <= (I.noff + I.nn - 1) / fsk::kWorkChunk; ++c)
        lim_h_.push_back(c * fsk::kWorkChunk);
    }
    std::sort(lim_h_.begin(), lim_h_.end());
    lim_h_.erase(std::unique(lim_h_.begin(), lim_h_.end()), lim_h_.end());
    if (!lim_h_.empty())
      FSE_HIP(hipMemcpy(d_lim_chunks_, lim_h_.data(), lim_h_.size() * sizeof(int), hipMemcpyHostToDevice));
    dm.lim_chunks = d_lim_chunks_;
    dm.n_lim_chunks = (int)lim_h_.size();
  }
  drop_graphs();
  return FS_OK;
}

int Engine::ensure_k2_fields(const std::vector<char>* want) {
  const int B = ninst();
  bool any = false;
  for (int i = 0; i < B; ++i) any = any || (k2_stale_[i] && (!want || (*want)[i]));
  if (!any) return FS_OK;
  // the staging buffer and the device table may still feed a step
  FSE_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < B; ++i) {
    const bool on = k2_stale_[i] && (!want || (*want)[i]);
    // the controls of the instance's last step (its counter has not moved
    // since), with the k=2 passes on
    const bool trial = k2_trial_;
    k2_trial_ = true;
    fsk::InstParams p = params_for(i, 1);
    k2_trial_ = trial;
    p.active = on ? 1 : 0;
    h_ip_[i] = p;
    if (on) note_phi(i, p.lim);
  }
  FSE_HIP(hipMemcpyAsync(d_ip_, h_ip_, (size_t)B * sizeof(fsk::InstParams), hipMemcpyHostToDevice, stream));
  // every node and tile (the work lists may no longer hold a finished instance)
  fsk::DevMesh m = dm;
  m.node_chunks = nullptr;
  m.n_node_chunks = 0;
  m.tile_list = nullptr;
  m.n_tile_list = 0;
  m.lim_chunks = nullptr;
  m.n_lim_chunks = 0;
  m.xcd_run = 0;
  fsk::DevState s = ds;
  s.nQ = qpre_;
  s.ip = d_ip_;
  fsk::launch_derivatives(m, s, stream);
  fsk::launch_limiter(m, s, stream);
  FSE_HIP(hipGetLastError());
  FSE_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < B; ++i)
    if (!want || (*want)[i]) k2_stale_[i] = 0;
  return FS_OK;
}

int Engine::upload_params(int it) {
  (void)it;
  // the staging buffer may still feed an earlier copy: drain the stream
  // (a no-op at a step boundary, where step_end has synchronised)
  FSE_HIP(hipStreamSynchronize(stream));
  const int B = ninst();
  for (int k = 1; k <= iter; ++k)
    for (int i = 0; i < B; ++i) h_ip_[(size_t)(k - 1) * B + i] = params_for(i, k);
  FSE_HIP(hipMemcpyAsync(d_ip_, h_ip_, (size_t)iter * B * sizeof(fsk::InstParams), hipMemcpyHostToDevice, stream));
  return FS_OK;
}

int Engine::upload_params_custom(const fsk::InstParams& p) {
  // a single pass may overwrite the deferred fields: evaluate them first
  if (int rc = ensure_k2_fields()) return rc;
  // and may write any phi
  std::fill(phi_one_.begin(), phi_one_.end(), 0);
  FSE_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < ninst(); ++i) h_ip_[i] = p;
  FSE_HIP(hipMemcpyAsync(d_ip_, h_ip_, (size_t)ninst() * sizeof(fsk::InstParams), hipMemcpyHostToDevice, stream));
  return FS_OK;
}

// Bit 0: some active instance needs the LSQ gradient; bit 1: k=2
// derivatives; -1: nothing is active.
int Engine::step_key() const {
  int key = 0;
  bool any = false;
  for (const auto& I : inst) {
    if (!I.active) continue;
    any = true;
    Phase ph = phase_for(I.prm, I.ts);
    if (ph.gradient) key |= 1;
    if (ph.k2 && !k2_deferred(I)) key |= 2;
    if (!(params_for((int)(&I - inst.data()), 1).lim & fsk::LIM_FROZEN)) key |= 4;
  }
  return any ? key : -1;
}

void Engine::drop_graphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

fsk::ColorRanges Engine::color_ranges() const {
  fsk::ColorRanges cr{color_off.data(), color_nback.data()};
  if (chunks_on_) {
    cr.chunk_off = chunk_off_h_.data();
    cr.chunks = d_chunks_;
  }
  return cr;
}

int Engine::refresh_work_lists() {
  const int B = ninst();
  long long act = 0;
  bool covered = chunks_on_;
  for (int i = 0; i < B; ++i) {
    if (!inst[i].active) continue;
    act += inst[i].nn;
    if (chunks_on_ && !chunk_inst_[i]) covered = false;
  }
  const bool want = act > 0 && 4 * act <= 3LL * dm.nn;
  if (!want) {
    if (chunks_on_) {
      chunks_on_ = false;
      dm.xcd_run = 0;
      dm.node_chunks = dm.edge_chunks = dm.pos_chunks = dm.tile_list = nullptr;
      dm.n_node_chunks = dm.n_edge_chunks = dm.n_pos_chunks = dm.n_tile_list = 0;
      drop_graphs();
    }
    return FS_OK;
  }
  if (covered && 4 * act > 3 * chunk_nodes_) return FS_OK;
  // an earlier list may still feed enqueued launches
  FSE_HIP(hipStreamSynchronize(stream));
  const int nc = (int)color_nback.size();
  constexpr int W = fsk::kWorkChunk;
  chunk_off_h_.assign(2 * (size_t)nc + 1, 0);
  chunk_h_.clear();
  // appends the first elements of the chunks (of `width` elements from
  // `origin`) covering [a, e); ranges arrive in increasing order
  auto add = [&](int origin, int width, int a, int e, size_t from) {
    if (e <= a) return;
    for (int k = (a - origin) / width; k <= (e - 1 - origin) / width; ++k) {
      const int first = origin + k * width;
      if (chunk_h_.size() == from || chunk_h_.back() < first) chunk_h_.push_back(first);
    }
  };
  // sweeps: per color, forward over the whole color and backward over its
  // first nback rows, in 16-row chunks from the range's start
  for (int c = 0; c < nc; ++c)
    for (int bwd = 0; bwd < 2; ++bwd) {
      const int b = color_off[c], e = bwd ? b + color_nback[c] : color_off[c + 1];
      const size_t from = chunk_h_.size();
      for (const Seg& g : color_segs_[c])
        if (inst[g.inst].active) add(b, fsk::kGsChunk, std::max(b, g.t0), std::min(e, g.t1), from);
      chunk_off_h_[2 * (size_t)c + bwd + 1] = (int)chunk_h_.size();
    }
  // nodes, edges, color positions: 64-element chunks from 0
  const size_t node0 = chunk_h_.size();
  for (int i = 0; i < B; ++i)
    if (inst[i].active) add(0, W, inst[i].noff, inst[i].noff + inst[i].nn, node0);
  const size_t edge0 = chunk_h_.size();
  for (int i = 0; i < B; ++i)
    if (inst[i].active) add(0, W, inst[i].eoff, inst[i].eoff + inst[i].ne, edge0);
  const size_t pos0 = chunk_h_.size();
  for (int c = 0; c < nc; ++c)
    for (const Seg& g : color_segs_[c])
      if (inst[g.inst].active) add(0, W, g.t0, g.t1, pos0);
  // LDS derivative tiles holding nodes of running k=2 instances
  const size_t tile0 = chunk_h_.size();
  if (dm.ndtile > 0)
    for (int i = 0; i < B; ++i) {
      const Instance& I = inst[i];
      if (!I.active || I.prm.order != 3 || I.nn == 0) continue;
      // the tiles containing the instance's first and last node
      auto tile_of = [&](int node) {
        return (int)(std::upper_bound(dt_off_h_.begin(), dt_off_h_.end(), node) - dt_off_h_.begin()) - 1;
      };
      const int a = tile_of(I.noff), z = tile_of(I.noff + I.nn - 1);
      for (int t = a; t <= z; ++t)
        if (chunk_h_.size() == tile0 || chunk_h_.back() < t) chunk_h_.push_back(t);
    }
  const size_t end = chunk_h_.size();
  if (end > 0) FSE_HIP(hipMemcpy(d_chunks_, chunk_h_.data(), end * sizeof(int), hipMemcpyHostToDevice));
  dm.node_chunks = d_chunks_ + node0;
  dm.n_node_chunks = (int)(edge0 - node0);
  dm.edge_chunks = d_chunks_ + edge0;
  dm.n_edge_chunks = (int)(pos0 - edge0);
  dm.pos_chunks = d_chunks_ + pos0;
  dm.n_pos_chunks = (int)(tile0 - pos0);
  dm.tile_list = d_chunks_ + tile0;
  dm.n_tile_list = (int)(end - tile0);
  chunk_inst_.assign(B, 0);
  for (int i = 0; i < B; ++i) chunk_inst_[i] = inst[i].active ? 1 : 0;
  chunk_nodes_ = act;
  chunks_on_ = true;
  // the running instances' nodes and rows are spread over all XCDs
  dm.xcd_run = fsk::kXcdRun;
  drop_graphs();
  return FS_OK;
}

// one Newton iteration of solve.C:328-909 over every active instance
int Engine::newton_iteration(int it, int key, bool record) {
  fsk::DevState s = ds;
  s.ip = d_ip_ + (size_t)(it - 1) * ninst();
  hipStream_t st = stream;
  if (key & 1) fsk::launch_gradient(dm, s, st);
  if (key & 2) fsk::launch_derivatives(dm, s, st);
  if (key & 4) fsk::launch_limiter(dm, s, st);
  const fsk::ColorRanges cr = color_ranges();
  if (precision != FS_PREC_F64) {
    // boundary fluxes and blocks, then one node pass for the FP64 residual,
    // ev_con/dt and every FP32 Jacobian block (and, fused, the sweep
    // operands: then the diagonal pass launches nothing); FP32 sweeps
    if (record) FSE_HIP(hipEventRecord(ev_[0], st));
    fsk::launch_bedge(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[1], st));
    if (record) FSE_HIP(hipEventRecord(ev_[2], st));
    fsk::launch_resid_jac32(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[3], st));
    fsk::launch_diag_lu32(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[5], st));
    fsk::launch_gs32_sweeps(dm, s, cr, subits, st);
    if (record) FSE_HIP(hipEventRecord(ev_[4], st));
    fsk::launch_update_mixed(dm, s, st);
    fsk::launch_norms(dm, s, nc, st);
  } else {
    if (record) FSE_HIP(hipEventRecord(ev_[0], st));
    fsk::launch_edge_flux(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[1], st));
    fsk::launch_bedge(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[2], st));
    fsk::launch_node(dm, s,
                     fsk::NODE_EVCON | fsk::NODE_INTERIOR | fsk::NODE_BOUNDARY | fsk::NODE_RHS | fsk::NODE_JAC, st);
    if (record) FSE_HIP(hipEventRecord(ev_[3], st));
    fsk::launch_diag_lu(dm, s, st);
    if (record) FSE_HIP(hipEventRecord(ev_[5], st));
    fsk::launch_gs_sweeps(dm, s, cr, subits, st);
    if (record) FSE_HIP(hipEventRecord(ev_[4], st));
    fsk::launch_norms(dm, s, nc, st);
    fsk::launch_update(dm, s, st);
  }
  FSE_HIP(hipGetLastError());
  return FS_OK;
}

int Engine::enqueue_step(int key, bool record) {
  // nQ = Q for real and ghost nodes (solve.C:319-325); instances deferring
  // their k=2 fields keep this state
  if (qpre_)
    fsk::launch_copy_state_save(dm, d_ip_, ds.nQ, ds.Q, qpre_, stream);
  else
    fsk::launch_copy_state(dm, ds.nQ, ds.Q, stream);
  for (int it = 1; it <= iter; ++it) {
    int rc = newton_iteration(it, key, record && it == iter);
    if (rc) return rc;
  }
  // Q = nQ for real and ghost nodes (solve.C:913-931)
  fsk::launch_copy_state(dm, ds.Q, ds.nQ, stream);
  FSE_HIP(hipGetLastError());
  return FS_OK;
}

int Engine::step(bool advance, bool graph, bool record) {
  if (record) {
    if (advance)
      for (auto& I : inst)
        if (I.active) ++I.ts;
    int key = step_key();
    if (key < 0) return FS_OK;
    int rc = refresh_work_lists();
    if (!rc) rc = upload_params(1);
    if (!rc) mark_k2_stale();
    if (!rc) rc = refresh_limiter_list();
    if (!rc) rc = enqueue_step(key, true);
    if (!rc) rc = enqueue_readback();
    if (rc) return rc;
    FSE_HIP(hipStreamSynchronize(stream));
    float a = 0, b = 0, c = 0, l = 0;
    FSE_HIP(hipEventElapsedTime(&a, ev_[0], ev_[1]));
    FSE_HIP(hipEventElapsedTime(&b, ev_[2], ev_[3]));
    FSE_HIP(hipEventElapsedTime(&l, ev_[3], ev_[5]));
    FSE_HIP(hipEventElapsedTime(&c, ev_[5], ev_[4]));
    t_edge += a; t_node += b; t_gs += c; t_lu += l;
    ++steps_run_;
    return check_errors(true);
  }
  int rc = step_begin(advance, graph);
  if (rc) return rc;
  return step_end();
}

int Engine::enqueue_readback() {
  const int B = ninst();
  FSE_HIP(hipMemcpyAsync(h_norms_, ds.inst_norms, 2 * (size_t)B * sizeof(double), hipMemcpyDeviceToHost, stream));
  FSE_HIP(hipMemcpyAsync(h_err_, ds.err, (size_t)B * sizeof(int), hipMemcpyDeviceToHost, stream));
  return FS_OK;
}

int Engine::step_begin(bool advance, bool graph) {
  if (pending_) {
    int rc = step_end();
    if (rc) return rc;
  }
  if (advance)
    for (auto& I : inst)
      if (I.active) ++I.ts;
  int key = step_key();
  if (key < 0) return FS_OK;
  int rc = refresh_work_lists();
  if (!rc) rc = upload_params(1);
  if (rc) return rc;
  mark_k2_stale();
  if ((rc = refresh_limiter_list())) return rc;
  if (graph) {
    auto itg = graphs_.find(key);
    if (itg == graphs_.end()) {
      hipGraph_t g = nullptr;
      FSE_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_step(key, false);
      hipError_t ce = hipStreamEndCapture(stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ce != hipSuccess) return set_error(FS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      hipGraphExec_t ex = nullptr;
      hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return set_error(FS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      itg = graphs_.emplace(key, ex).first;
    }
    FSE_HIP(hipGraphLaunch(itg->second, stream));
  } else {
    rc = enqueue_step(key, false);
    if (rc) return rc;
  }
  rc = enqueue_readback();
  if (rc) return rc;
  pending_ = true;
  return FS_OK;
}

int Engine::step_end() {
  if (!pending_) return FS_OK;
  pending_ = false;
  ++steps_run_;
  FSE_HIP(hipStreamSynchronize(stream));
  return check_errors(true);
}

// Compute_Body_Forces (grid_io.C:1390-1697) of every instance from its
// resident state; citer[i] (grid->citer) selects the boundary reconstruction
// as Reconstruct_Gauss_Boundary does (reconstruct.C:1733-1770).  The device
// evaluates the per-bedge terms; their sum runs here in the reference's
// bedge order over the instance's walls (odd bc), so the result is the
// reference's sequential sum.
int Engine::body_forces(const int* citer, double* forces, const std::vector<char>* want) {
  if (int rc = ensure_k2_fields(want)) return rc;
  const int B = ninst();
  for (int i = 0; i < B; ++i) {
    const fs_params& p = inst[i].prm;
    const Phase ph = phase_for(p, citer[i]);
    const double angle = p.alpha * M_PI / 180;
    h_fp_[i].rorder = ph.k2 ? 3 : ph.recon_flux;
    h_fp_[i].gamma = p.gamma;
    h_fp_[i].ca = std::cos(angle);
    h_fp_[i].sa = std::sin(angle);
  }
  FSE_HIP(hipMemcpyAsync(d_fp_, h_fp_, (size_t)B * sizeof(fsk::ForceParams), hipMemcpyHostToDevice, stream));
  fsk::launch_body_terms(dm, ds, d_fp_, stream);
  FSE_HIP(hipGetLastError());
  if (dm.nbe > 0)
    FSE_HIP(hipMemcpyAsync(h_bforce_, ds.bforce, 2 * (size_t)dm.nbe * sizeof(double), hipMemcpyDeviceToHost, stream));
  FSE_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < B; ++i) {
    const Instance& I = inst[i];
    const fs::Mesh& g = *I.mesh;
    double f0 = 0., f1 = 0.;
    for (int b = 0; b < I.nbe; ++b) {
      const int bc = g.bbc[g.bedge[5 * (size_t)b + 3]];
      if (bc == 0 || (bc % 2 == 0)) continue;
      f0 += h_bforce_[2 * (size_t)(I.boff + b)];
      f1 += h_bforce_[2 * (size_t)(I.boff + b) + 1];
    }
    forces[2 * (size_t)i] = f0;
    forces[2 * (size_t)i + 1] = f1;
  }
  return FS_OK;
}

int Engine::fetch_errors() {
  FSE_HIP(hipMemcpyAsync(h_err_, ds.err, (size_t)ninst() * sizeof(int), hipMemcpyDeviceToHost, stream));
  FSE_HIP(hipStreamSynchronize(stream));
  return check_errors(false);
}

int Engine::check_errors(bool norms) {
  bool any = false;
  for (int i = 0; i < ninst(); ++i) {
    Instance& I = inst[i];
    if (!I.active) continue;
    if (norms) {
      I.r2 = h_norms_[2 * (size_t)i];
      I.dq2 = h_norms_[2 * (size_t)i + 1];
    }
    I.err = h_err_[i];
    if (I.err) any = true;
  }
  if (any) {
    std::memset(h_err_, 0, (size_t)ninst() * sizeof(int));
    FSE_HIP(hipMemsetAsync(ds.err, 0, (size_t)ninst() * sizeof(int), stream));
    FSE_HIP(hipStreamSynchronize(stream));
  }
  return FS_OK;
}

int Engine::sync() {
  FSE_HIP(hipStreamSynchronize(stream));
  return FS_OK;
}

// ---------------------------------------------------------------------------
// recovery checkpoints
// ---------------------------------------------------------------------------
int Engine::enable_checkpoints() {
  if (qsave_) return FS_OK;
  if (dm.ng < dm.nbe) return set_error(FS_ERR_STATE, "checkpoints: fewer ghosts than boundary edges");
  int rc = alloc(&qsave_, 4 * ((size_t)dm.nn + dm.ng));
  if (!rc) rc = alloc(&qphase_, 4 * ((size_t)dm.nn + dm.ng));
  if (!rc) rc = alloc(&d_rmode_, (size_t)ninst());
  if (rc) return rc;
  FSE_HIP(hipHostMalloc((void**)&h_rmode_, (size_t)ninst() * sizeof(int)));
  for (auto& I : inst) I.saved_ts = I.phase_ts = -1;   // the first healthy step saves
  return FS_OK;
}

int Engine::checkpoint_pass(const std::vector<int>& mode) {
  if (!qsave_) return set_error(FS_ERR_STATE, "checkpoints are not enabled");
  if ((int)mode.size() != ninst()) return set_error(FS_ERR_ARG, "checkpoint modes: one per instance");
  bool any = false;
  for (int m : mode) any = any || m != 0;
  if (!any) return FS_OK;
  // the staging buffer may still feed the previous pass's copy
  FSE_HIP(hipStreamSynchronize(stream));
  std::copy(mode.begin(), mode.end(), h_rmode_);
  FSE_HIP(hipMemcpyAsync(d_rmode_, h_rmode_, (size_t)ninst() * sizeof(int), hipMemcpyHostToDevice, stream));
  fsk::launch_checkpoint(dm, ds.Q, qsave_, qphase_, d_rmode_, stream);
  FSE_HIP(hipGetLastError());
  return FS_OK;
}

namespace {
int phase_id(const fs_params& p, int ts) {
  const Phase ph = phase_for(p, ts);
  return ph.recon_flux | (ph.k2 ? 4 : 0) | (ph.lim << 3);
}
}  // namespace

bool Engine::recover_or_checkpoint(int i, bool failed, double rn, const RecoveryPolicy& pol, RecoveryTrack& t,
                                   std::vector<int>& mode) {
  if (pol.max_recoveries <= 0 || !qsave_) return false;
  Instance& I = inst[i];
  // the fallbacks restart the current phase from its start (the phase
  // checkpoint holds the state the phase began from); returns the step
  // restored to, or -1 when none applies
  bool cfl_reset = false;
  auto fallback = [&]() -> int {
    const bool at_phase = I.phase_ts >= 0 && phase_id(I.prm, I.phase_ts + 1) == phase_id(I.prm, I.ts);
    if (!at_phase) return -1;
    const Phase ph = phase_for(I.prm, I.ts);
    // the limiter fallback: an unlimited solve in a phase that would limit
    fs_params lim = I.prm;
    lim.limit = 1;
    if (pol.limiter_fallback > 0 && I.prm.limit == 0 && (phase_for(lim, I.ts).lim & fsk::LIM_ON)) {
      I.prm.limit = 1;
      I.freeze = pol.limiter_fallback;
      t.fallbacks |= kFallbackLimiter;
    } else if (pol.first_order_fallback && !(t.fallbacks & kFallbackFirstOrder) && (ph.recon_flux == 1 || ph.k2)) {
      // the first-order start-up phase never ends: nodal face states, no
      // gradient, derivatives or limiter from here on; the CFL target the
      // restores lowered returns to the solve's own (the first-order
      // iteration is the robust one, and at a quarter of the target its
      // convergence is linear and slow)
      I.prm.foits = 1 << 28;
      I.cfl_scale = 1.0;
      cfl_reset = true;
      t.fallbacks |= kFallbackFirstOrder;
    } else {
      return -1;
    }
    mode[i] = fsk::CKPT_RESTORE_PHASE;
    I.saved_ts = -1;   // the latest checkpoint is later than the restart
    return I.phase_ts;
  };
  auto restart = [&](int to) {
    ++t.recoveries;
    I.err = 0;
    I.ts = to;
    I.ramp_from = to;
    t.since_save = 0;
    t.phase = -1;
  };
  if (failed) {
    if (t.recoveries >= pol.max_recoveries) return false;
    int to = fallback();
    if (to < 0) {
      if (I.saved_ts >= 0) {
        mode[i] = fsk::CKPT_RESTORE;
        to = I.saved_ts;
      } else if (I.phase_ts >= 0) {
        mode[i] = fsk::CKPT_RESTORE_PHASE;
        to = I.phase_ts;
      } else {
        return false;
      }
    }
    restart(to);
    if (!cfl_reset) I.cfl_scale *= 0.5;
    return true;
  }
  const int pid = phase_id(I.prm, I.ts);
  if (pid != t.phase) {
    t.phase = pid;
    t.r_best = rn;
    t.r_mark = rn;
    t.since_mark = 0;
  } else if (rn < 0.5 * t.r_mark) {
    t.r_mark = rn;
    t.since_mark = 0;
  } else {
    ++t.since_mark;
  }
  if (pol.stall_steps_per_sqrt_node > 0 && t.recoveries < pol.max_recoveries &&
      t.since_mark >= std::max(kStallMinSteps, (int)std::ceil(pol.stall_steps_per_sqrt_node * std::sqrt((double)I.nn)))) {
    const int to = fallback();
    if (to >= 0) {
      t.fallbacks |= kFallbackStall;
      restart(to);
      return true;
    }
  }
  ++t.since_save;
  if ((I.saved_ts < 0 || t.since_save >= std::max(1, pol.every)) && rn <= 10.0 * t.r_best) {
    mode[i] |= fsk::CKPT_SAVE;
    I.saved_ts = I.ts;
    t.since_save = 0;
  }
  // the last step of a phase: the state the next phase starts from
  if (phase_id(I.prm, I.ts + 1) != pid) {
    mode[i] |= fsk::CKPT_SAVE_PHASE;
    I.phase_ts = I.ts;
  }
  t.r_best = std::min(t.r_best, rn);
  return false;
}

int Engine::adopt_checkpoint(int i, Engine& from, int j) {
  if (!qsave_ || !from.qsave_) return set_error(FS_ERR_STATE, "checkpoints are not enabled");
  const Instance& F = from.inst[j];
  Instance& T = inst[i];
  for (int slot = 0; slot < 2; ++slot) {
    double* dst = slot ? qphase_ : qsave_;
    const double* src = slot ? from.qphase_ : from.qsave_;
    FSE_HIP(hipMemcpyAsync(dst + 4 * (size_t)T.noff, src + 4 * (size_t)F.noff, 4 * (size_t)F.nn * sizeof(double),
                           hipMemcpyDeviceToDevice, stream));
    if (F.nbe > 0)
      FSE_HIP(hipMemcpyAsync(dst + 4 * ((size_t)dm.nn + T.boff), src + 4 * ((size_t)from.dm.nn + F.boff),
                             4 * (size_t)F.nbe * sizeof(double), hipMemcpyDeviceToDevice, stream));
  }
  T.saved_ts = F.saved_ts;
  T.phase_ts = F.phase_ts;
  T.cfl_scale = F.cfl_scale;
  T.ramp_from = F.ramp_from;
  T.prm.limit = F.prm.limit;
  T.prm.foits = F.prm.foits;
  T.freeze = F.freeze;
  return FS_OK;
}

// ---------------------------------------------------------------------------
// autotuning of launch shapes
// ---------------------------------------------------------------------------
std::string Engine::tune_key() const {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    arch = prop.gcnArchName;
    size_t c = arch.find(':');
    if (c != std::string::npos) arch = arch.substr(0, c);
  }
  bool k2 = false;
  for (const auto& I : inst) k2 = k2 || I.prm.order == 3;
  int lg = 0;
  while ((1LL << (lg + 1)) <= (long long)std::max(1, dm.nn)) ++lg;
  std::ostringstream k;
  k << arch << ":" << (precision == FS_PREC_MIXED ? "mixed" : precision == FS_PREC_MIXED16 ? "mixed16" : "f64") << ":"
    << (k2 ? "k2" : "k1") << ":n" << lg;
  return k.str();
}

int Engine::set_tuning(const int* v) {
  fsk::Tuning t;
  t.gs_block = v[0]; t.gs_lanes = v[1]; t.gs32_block = v[2]; t.gs32_pre = v[3];
  t.node_block = v[4]; t.face_block = v[5]; t.cell_block = v[6]; t.deriv_lds = v[7]; t.gs32_scaled = v[8];
  t.node_fused = v[9];
  t.gs32_dpp = v[10];
  if (!valid_tuning(t)) return set_error(FS_ERR_ARG, "tuning value outside the candidates");
  if (precision == FS_PREC_MIXED16 && !t.gs32_scaled)
    return set_error(FS_ERR_ARG, "FP16 sweep operands exist only pre-multiplied (gs32_scaled 1)");
  FSE_HIP(hipStreamSynchronize(stream));
  dm.tune = t;
  // captured graphs hold the old launch shapes
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  return FS_OK;
}

int Engine::autotune(double* seconds) {
  if (steps_run_ < 1) return set_error(FS_ERR_STATE, "autotune needs one completed step (run a step first)");
  auto t_start = std::chrono::steady_clock::now();
  int rc = sync();
  if (rc) return rc;
  // the trials run the k=2 passes as a step that does not defer them would
  // (they overwrite the fields; deferred ones are still evaluated later from
  // the kept state)
  k2_trial_ = true;
  rc = upload_params(1);
  const int key = step_key() < 0 ? 0 : step_key();
  k2_trial_ = false;
  if (rc) return rc;
  for (int i = 0; i < ninst(); ++i)
    if (inst[i].active) note_phi(i, h_ip_[i].lim);   // the trials run the limiter pass
  // the limiter trials cover the nodes the step's work lists cover; the next
  // step rebuilds the limiter's own list
  dm.lim_chunks = nullptr;
  dm.n_lim_chunks = 0;
  lim_set_.assign(ninst(), 2);
  fsk::DevState s = ds;  // controls of the last step's first Newton iteration
  fsk::ColorRanges cr{color_off.data(), color_nback.data()};
  const int blocks[] = {64, 128, 256, 512};
  hipStream_t st = stream;
  // mean device time of `reps` launches of `f` after one untimed launch
  auto time_it = [&](auto&& f, int reps, float* ms) -> int {
    f();
    FSE_HIP(hipEventRecord(ev_[6], st));
    for (int r = 0; r < reps; ++r) f();
    FSE_HIP(hipEventRecord(ev_[7], st));
    FSE_HIP(hipEventSynchronize(ev_[7]));
    FSE_HIP(hipGetLastError());
    float t = 0;
    FSE_HIP(hipEventElapsedTime(&t, ev_[6], ev_[7]));
    *ms = t / reps;
    return FS_OK;
  };
  fsk::Tuning& tu = dm.tune;
  float best, t;
  if (precision != FS_PREC_F64) {
    best = 1e30f;
    fsk::Tuning keep = tu;
    for (int dpp : {0, 1})
      for (int pre : {4, 6, 8})
        for (int b : blocks) {
          tu.gs32_dpp = dpp;
          tu.gs32_pre = pre;
          tu.gs32_block = b;
          if ((rc = time_it([&] { fsk::launch_gs32_sweeps(dm, s, cr, 2, st); }, 2, &t))) return rc;
          if (t < best) { best = t; keep.gs32_pre = pre; keep.gs32_block = b; keep.gs32_dpp = dpp; }
        }
    tu.gs32_pre = keep.gs32_pre;
    tu.gs32_block = keep.gs32_block;
    tu.gs32_dpp = keep.gs32_dpp;
  } else {
    best = 1e30f;
    fsk::Tuning keep = tu;
    // 1 thread per row, or 4 lanes per row with each prefetch depth and gather form
    const int forms[][3] = {{1, 6, 0}, {4, 0, 0}, {4, 4, 0}, {4, 6, 0}, {4, 8, 0}, {4, 4, 1}, {4, 6, 1}, {4, 8, 1}};
    for (const auto& f : forms)
      for (int b : blocks) {
        tu.gs_lanes = f[0];
        tu.gs32_pre = f[1];
        tu.gs32_dpp = f[2];
        tu.gs_block = b;
        if ((rc = time_it([&] { fsk::launch_gs_sweeps(dm, s, cr, 2, st); }, 2, &t))) return rc;
        if (t < best) {
          best = t;
          keep.gs_lanes = f[0]; keep.gs32_pre = f[1]; keep.gs32_dpp = f[2]; keep.gs_block = b;
        }
      }
    tu.gs_lanes = keep.gs_lanes;
    tu.gs_block = keep.gs_block;
    tu.gs32_pre = keep.gs32_pre;
    tu.gs32_dpp = keep.gs32_dpp;
  }
  const int nflags = fsk::NODE_EVCON | fsk::NODE_INTERIOR | fsk::NODE_BOUNDARY | fsk::NODE_RHS | fsk::NODE_JAC;
  int keepb = tu.node_block;
  best = 1e30f;
  // the mixed node pass with the diagonal pass (a no-op when the node pass
  // forms the sweep operands itself), which always starts from the blocks
  // the node pass has just written
  auto node_pass32 = [&] {
    fsk::launch_resid_jac32(dm, s, st);
    fsk::launch_diag_lu32(dm, s, st);
  };
  for (int b : blocks) {
    tu.node_block = b;
    if (precision != FS_PREC_F64)
      rc = time_it(node_pass32, 2, &t);
    else
      rc = time_it([&] { fsk::launch_node(dm, s, nflags, st); }, 2, &t);
    if (rc) return rc;
    if (t < best) { best = t; keepb = b; }
  }
  tu.node_block = keepb;
  if (precision == FS_PREC_MIXED && tu.gs32_scaled) {
    // pre-multiplied operands from the node pass itself or from the diagonal pass
    int keepf = tu.node_fused;
    best = 1e30f;
    for (int f : {0, 1}) {
      tu.node_fused = f;
      if ((rc = time_it(node_pass32, 2, &t))) return rc;
      if (t < best) { best = t; keepf = f; }
    }
    tu.node_fused = keepf;
  }
  if (precision == FS_PREC_F64) {
    keepb = tu.face_block;
    best = 1e30f;
    for (int b : blocks) {
      tu.face_block = b;
      rc = time_it([&] { fsk::launch_edge_flux(dm, s, st); }, 2, &t);
      if (rc) return rc;
      if (t < best) { best = t; keepb = b; }
    }
    tu.face_block = keepb;
  }
  keepb = tu.cell_block;
  best = 1e30f;
  // the diagonal pass is timed in its inverse-only form: the pre-multiplied
  // form would scale the blocks again at every trial
  const int scaled = tu.gs32_scaled;
  tu.gs32_scaled = 0;
  for (int b : blocks) {
    tu.cell_block = b;
    rc = time_it([&] {
      if (key & 1) fsk::launch_gradient(dm, s, st);
      if (key & 2) fsk::launch_derivatives(dm, s, st);
      fsk::launch_limiter(dm, s, st);
      if (precision == FS_PREC_MIXED) fsk::launch_diag_lu32(dm, s, st);
      else if (precision == FS_PREC_F64) fsk::launch_diag_lu(dm, s, st);
    }, 2, &t);
    if (rc) {
      tu.gs32_scaled = scaled;
      return rc;
    }
    if (t < best) { best = t; keepb = b; }
  }
  tu.gs32_scaled = scaled;
  tu.cell_block = keepb;
  if ((key & 2) && dm.ndtile > 0) {
    // the k=2 derivatives: neighbour states gathered through L2 or staged in LDS
    int keepl = tu.deriv_lds;
    best = 1e30f;
    for (int l : {0, 1}) {
      tu.deriv_lds = l;
      if ((rc = time_it([&] { fsk::launch_derivatives(dm, s, st); }, 3, &t))) return rc;
      if (t < best) { best = t; keepl = l; }
    }
    tu.deriv_lds = keepl;
  }
  // captured graphs hold the old launch shapes
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  // record the choice
  std::string path = tune_cache_path();
  if (!path.empty()) {
    size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
      std::string dir = path.substr(0, slash);
      for (size_t i = 1; i <= dir.size(); ++i)
        if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    }
    std::ofstream out(path, std::ios::app);
    if (out)
      out << tune_key() << " " << tu.gs_block << " " << tu.gs_lanes << " " << tu.gs32_block << " " << tu.gs32_pre << " "
          << tu.node_block << " " << tu.face_block << " " << tu.cell_block << " " << tu.deriv_lds << " "
          << tu.node_fused << " " << tu.gs32_dpp << "\n";
  }
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  return FS_OK;
}

// ---------------------------------------------------------------------------
// per-instance fields
// ---------------------------------------------------------------------------
int Engine::field_count(int i, int field, int64_t* n) const {
  const Instance& I = inst[i];
  int64_t nn = I.nn, ntot = (int64_t)I.nn + I.nbe;
  switch (field) {
    case FS_F_Q: case FS_F_NQ: *n = ntot * 4; return FS_OK;
    case FS_F_DQ: case FS_F_R: case FS_F_RHS: case FS_F_PV: *n = nn * 4; return FS_OK;
    case FS_F_GRAD: *n = nn * 8; return FS_OK;
    case FS_F_PHI: case FS_F_EVCON: case FS_F_DT: *n = nn; return FS_OK;
    case FS_F_HESS: *n = nn * 20; return FS_OK;
    case FS_F_LHS: case FS_F_LHS32: *n = (int64_t)I.nnz * 16; return FS_OK;
    default: return set_error(FS_ERR_ARG, "unknown field id");
  }
}

static double* node_field(const fsk::DevState& s, int field, int* per) {
  switch (field) {
    case FS_F_Q: *per = 4; return s.Q;
    case FS_F_NQ: *per = 4; return s.nQ;
    case FS_F_DQ: *per = 4; return s.dQ;
    case FS_F_R: *per = 4; return s.R;
    case FS_F_RHS: *per = 4; return s.RHS;
    case FS_F_PV: *per = 4; return s.pv;
    case FS_F_GRAD: *per = 8; return s.grad;
    case FS_F_PHI: *per = 1; return s.phi;
    case FS_F_EVCON: *per = 1; return s.ev_con;
    case FS_F_DT: *per = 1; return s.dt;
    case FS_F_HESS: *per = 20; return s.hess;
    default: *per = 0; return nullptr;
  }
}

int Engine::get_field(int i, int field, double* host, int64_t n) {
  if (i < 0 || i >= ninst()) return set_error(FS_ERR_ARG, "instance out of range");
  int64_t cnt = 0;
  int rc = field_count(i, field, &cnt);
  if (rc) return rc;
  if (n != cnt) return set_error(FS_ERR_ARG, "size mismatch: expected " + std::to_string(cnt));
  if (field == FS_F_PHI || field == FS_F_HESS || field == FS_F_PV) {
    std::vector<char> one(ninst(), 0);
    one[i] = 1;
    if ((rc = ensure_k2_fields(&one))) return rc;
  }
  const Instance& I = inst[i];
  FSE_HIP(hipStreamSynchronize(stream));
  if (field == FS_F_LHS || field == FS_F_LHS32) {
    if (field == FS_F_LHS && !ds.LHS) return set_error(FS_ERR_ARG, "FP64 blocks exist only for an FP64 solver");
    if (field == FS_F_LHS32 && !ds.D32) return set_error(FS_ERR_ARG, "FP32 blocks exist only for a mixed-precision solver");
    size_t all = (size_t)dm.nnz * 16;
    if (field == FS_F_LHS) {
      std::vector<double> tmp(all);
      FSE_HIP(hipMemcpy(tmp.data(), ds.LHS, all * sizeof(double), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < I.slot_dev.size(); ++k)
        std::memcpy(host + 16 * k, &tmp[16 * (size_t)I.slot_dev[k]], 16 * sizeof(double));
    } else if (ds.LHS16) {
      // FP16 off-diagonal operands widened, FP32 diagonal blocks
      const size_t noff = (size_t)dm.nnz - dm.nn;
      std::vector<_Float16> h(noff * 16);
      std::vector<float> dg((size_t)dm.nn * 16);
      FSE_HIP(hipMemcpy(h.data(), ds.LHS16, h.size() * sizeof(_Float16), hipMemcpyDeviceToHost));
      FSE_HIP(hipMemcpy(dg.data(), ds.D32, dg.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < I.slot_dev.size(); ++k) {
        const size_t sl = (size_t)I.slot_dev[k];
        for (int c = 0; c < 16; ++c)
          host[16 * k + c] = sl < noff ? (double)(float)h[16 * sl + c] : (double)dg[16 * (sl - noff) + c];
      }
    } else {
      std::vector<float> tmp(all);
      FSE_HIP(hipMemcpy(tmp.data(), ds.LHS32, all * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < I.slot_dev.size(); ++k)
        for (int c = 0; c < 16; ++c) host[16 * k + c] = (double)tmp[16 * (size_t)I.slot_dev[k] + c];
    }
    return FS_OK;
  }
  int per = 0;
  double* d = node_field(ds, field, &per);
  FSE_HIP(hipMemcpy(host, d + (size_t)per * I.noff, (size_t)per * I.nn * sizeof(double), hipMemcpyDeviceToHost));
  if (field == FS_F_Q || field == FS_F_NQ)
    FSE_HIP(hipMemcpy(host + (size_t)per * I.nn, d + (size_t)per * ((size_t)dm.nn + I.boff),
                      (size_t)per * I.nbe * sizeof(double), hipMemcpyDeviceToHost));
  return FS_OK;
}

int Engine::set_field(int i, int field, const double* host, int64_t n) {
  if (i < 0 || i >= ninst()) return set_error(FS_ERR_ARG, "instance out of range");
  int64_t cnt = 0;
  int rc = field_count(i, field, &cnt);
  if (rc) return rc;
  if (n != cnt) return set_error(FS_ERR_ARG, "size mismatch: expected " + std::to_string(cnt));
  {
    // the instance's deferred fields come from the state before this write
    std::vector<char> one(ninst(), 0);
    one[i] = 1;
    if ((rc = ensure_k2_fields(&one))) return rc;
  }
  if (field == FS_F_PHI) phi_one_[i] = 0;
  const Instance& I = inst[i];
  FSE_HIP(hipStreamSynchronize(stream));
  if (field == FS_F_LHS32) return set_error(FS_ERR_ARG, "FP32 blocks are read-only");
  if (field == FS_F_LHS) {
    if (!ds.LHS) return set_error(FS_ERR_ARG, "FP64 blocks exist only for an FP64 solver");
    size_t all = (size_t)dm.nnz * 16;
    std::vector<double> tmp(all);
    FSE_HIP(hipMemcpy(tmp.data(), ds.LHS, all * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < I.slot_dev.size(); ++k)
      std::memcpy(&tmp[16 * (size_t)I.slot_dev[k]], host + 16 * k, 16 * sizeof(double));
    FSE_HIP(hipMemcpy(ds.LHS, tmp.data(), all * sizeof(double), hipMemcpyHostToDevice));
    return FS_OK;
  }
  int per = 0;
  double* d = node_field(ds, field, &per);
  FSE_HIP(hipMemcpy(d + (size_t)per * I.noff, host, (size_t)per * I.nn * sizeof(double), hipMemcpyHostToDevice));
  if (field == FS_F_Q || field == FS_F_NQ)
    FSE_HIP(hipMemcpy(d + (size_t)per * ((size_t)dm.nn + I.boff), host + (size_t)per * I.nn,
                      (size_t)per * I.nbe * sizeof(double), hipMemcpyHostToDevice));
  return FS_OK;
}

}  // namespace fse
