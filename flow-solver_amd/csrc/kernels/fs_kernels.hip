// FP64 kernels of the residual-and-implicit-update path (gfx950, wave64).
//
// Deterministic by construction: the per-node sums of the residual, the
// gradient and ev_con are gathers over the node's incident subedges/edges/
// bedges in ascending index order, which is the order in which the
// reference's sequential scatter loops add into a node (residual.C:112-333
// then residual.C:429-1193; gradient.C:382-403; cv_calc.C:819-953).  Face
// fluxes are computed once per subedge by an edge-centric kernel and written
// to a per-incidence buffer; the node kernel reads them back.  Other sums
// (k=2 derivatives, norms) use fixed lane partitions and butterflies.  No
// floating-point atomics anywhere.
//
// Every element reads its instance's controls (InstParams) and does nothing
// for an inactive instance, so one launch serves a packed batch of meshes and
// an instance's arithmetic never depends on the other instances.
#include "fs_device.hpp"
#include "fs_kernels.hpp"

#include <vector>

namespace fsk {
using namespace fsd;

namespace {
constexpr int kBlock = 256;
inline int nblocks(int n, int b = kBlock) { return (n + b - 1) / b; }
__device__ __forceinline__ double dmin(double x, double y) { return x <= y ? x : y; }
__device__ __forceinline__ double dmax(double x, double y) { return x >= y ? x : y; }
}  // namespace

// ---------------------------------------------------------------------------
// K2: least-squares gradient, node-centric gather (gradient.C:363-403)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxBlock) k_gradient(DevMesh m, DevState s) {
  int n = work_elem(m.node_chunks, m.n_node_chunks, xcd_block(m.xcd_run) * blockDim.x + threadIdx.x, m.nn);
  if (n >= m.nn) return;
  const InstParams& ip = s.ip[inst_of(m, m.inst_node, n)];
  if (!ip.active || !ip.grad) return;
  const double* __restrict__ nQ = s.nQ;
  double* __restrict__ grad = s.grad;
  double g[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) g[k] = 0.0;
  double qn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) qn[j] = nQ[4 * (size_t)n + j];
  for (int k = m.nedge_off[n]; k < m.nedge_off[n + 1]; ++k) {
    int ent = m.inc_other[k];
    int o = ent & 0x7fffffff;
    double w0 = m.inc_gw[2 * (size_t)k], w1 = m.inc_gw[2 * (size_t)k + 1];
    if (ent >= 0) {  // this node is L: dQ = Q_R - Q_L, grad_L += w dQ
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double dq = nQ[4 * (size_t)o + j] - qn[j];
        g[2 * j] += w0 * dq;
        g[2 * j + 1] += w1 * dq;
      }
    } else {         // this node is R: grad_R -= w dQ
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double dq = qn[j] - nQ[4 * (size_t)o + j];
        g[2 * j] -= w0 * dq;
        g[2 * j + 1] -= w1 * dq;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) grad[8 * (size_t)n + k] = g[k];
}

// ---------------------------------------------------------------------------
// K3: k=2 derivatives, sol = W' (q_j - q_i) over the 2-ring stencil
// (compute_derivatives.C:1438-1521).  16 lanes per node: lane v + 4p takes
// variable v over the stencil entries p, p+4, ...; the 4 partial sums of
// each moment are combined with a fixed butterfly (xor 4, then xor 8).  W' is
// stored entry-major ([entry][5]), so the lanes of a node read 4 consecutive
// entries per step and the 4 variable lanes of an entry share its cache line.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxBlock) k_derivatives(DevMesh m, DevState s) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  int i = work_elem(m.node_chunks, m.n_node_chunks, tid >> 4, m.nn), v = tid & 3, p = (tid >> 2) & 3;
  if (i >= m.nn) return;
  const InstParams& ip = s.ip[inst_of(m, m.inst_node, i)];
  if (!ip.active || !ip.k2) return;
  const double* __restrict__ nQ = s.nQ;
  const int* __restrict__ st_nodes = m.st_nodes;
  const double* __restrict__ Wt = m.st_Wt;
  const int o = m.st_off[i], e = m.st_off[i + 1];
  const double qi = nQ[4 * (size_t)i + v];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
  // the first kHead entries of this lane: all index and state loads are
  // issued before the first product, then the rest of a long stencil
  constexpr int kHead = 5;
  int nd[kHead];
#pragma unroll
  for (int k = 0; k < kHead; ++k) {
    int j = o + p + 4 * k;
    nd[k] = (j < e) ? st_nodes[j] : -1;
  }
  double d[kHead];
#pragma unroll
  for (int k = 0; k < kHead; ++k) d[k] = (nd[k] >= 0) ? nQ[4 * (size_t)nd[k] + v] - qi : 0.0;
#pragma unroll
  for (int k = 0; k < kHead; ++k) {
    if (nd[k] < 0) break;
    const double* w = Wt + 5 * (size_t)(o + p + 4 * k);
    a0 += w[0] * d[k];
    a1 += w[1] * d[k];
    a2 += w[2] * d[k];
    a3 += w[3] * d[k];
    a4 += w[4] * d[k];
  }
  for (int j = o + p + 4 * kHead; j < e; j += 4) {
    const double dj = nQ[4 * (size_t)st_nodes[j] + v] - qi;
    const double* w = Wt + 5 * (size_t)j;
    a0 += w[0] * dj;
    a1 += w[1] * dj;
    a2 += w[2] * dj;
    a3 += w[3] * dj;
    a4 += w[4] * dj;
  }
#pragma unroll
  for (int off = 4; off <= 8; off <<= 1) {
    a0 += __shfl_xor(a0, off, 16);
    a1 += __shfl_xor(a1, off, 16);
    a2 += __shfl_xor(a2, off, 16);
    a3 += __shfl_xor(a3, off, 16);
    a4 += __shfl_xor(a4, off, 16);
  }
  if (p != 0) return;
  double* __restrict__ h = s.hess + 20 * (size_t)i + 5 * v;
  h[0] = a0;
  h[1] = a1;
  h[2] = a2;
  h[3] = a3;
  h[4] = a4;
  const double* Mi = m.moments + 9 * (size_t)i;
  s.pv[4 * (size_t)i + v] = qi - Mi[0] * a0 - Mi[1] * a1 - Mi[2] * a2 * 0.5 - Mi[3] * a3 * 0.5 - Mi[4] * a4;
}

// Lane r's double in every lane of its quad (DPP quad_perm [r,r,r,r] on
// both halves).  Every lane of the quad must be active.
template <int kLane>
__device__ __forceinline__ double quad_bcast_d(double x) {
  constexpr int ctrl = kLane * 0x55;
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The same pass with a tile's neighbour states staged in LDS: the block of
// tile t (kDerivTileNodes nodes, 16 lanes each) first copies the states of
// the tile's distinct stencil nodes into LDS, then every lane runs the loop
// above reading them by slot.  Each lane sums the same entries in the same
// order, so the results are those of k_derivatives bit for bit.
__global__ void __launch_bounds__(16 * kDerivTileNodes) k_derivatives_lds(DevMesh m, DevState s) {
  __shared__ double sq[4 * kDerivTileMaxU];
  int t = xcd_block(m.xcd_run);
  if (m.tile_list) t = m.tile_list[t];   // one workgroup per listed tile
  const double* __restrict__ nQ = s.nQ;
  const int i = m.dt_off[t] + (threadIdx.x >> 4), v = threadIdx.x & 3, p = (threadIdx.x >> 2) & 3;
  bool mine = false;
  if (i < m.dt_off[t + 1]) {
    const InstParams& ip = s.ip[inst_of(m, m.inst_node, i)];
    mine = ip.active && ip.k2;
  }
  // a tile none of whose nodes needs derivatives this iteration stages nothing
  if (!__syncthreads_or(mine)) return;
  const int u0 = m.dt_uoff[t], nu = m.dt_uoff[t + 1] - u0;
  for (int k = threadIdx.x; k < 4 * nu; k += blockDim.x) sq[k] = nQ[4 * (size_t)m.dt_unodes[u0 + (k >> 2)] + (k & 3)];
  __syncthreads();
  if (!mine) return;
  const unsigned short* __restrict__ loc = m.st_loc;
  const double* __restrict__ Wt = m.st_Wt;
  const int o = m.st_off[i], e = m.st_off[i + 1];
  const double qi = nQ[4 * (size_t)i + v];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
  // the 4 lanes of a quad (variables 0-3) share stencil entry j: lane v loads
  // its weight v and the quad exchanges them (weight 4 is loaded by all)
  for (int j = o + p; j < e; j += 4) {
    const double wv = Wt[5 * (size_t)j + v], w4 = Wt[5 * (size_t)j + 4];
    const double dj = sq[4 * loc[j] + v] - qi;
    a0 += quad_bcast_d<0>(wv) * dj;
    a1 += quad_bcast_d<1>(wv) * dj;
    a2 += quad_bcast_d<2>(wv) * dj;
    a3 += quad_bcast_d<3>(wv) * dj;
    a4 += w4 * dj;
  }
#pragma unroll
  for (int off = 4; off <= 8; off <<= 1) {
    a0 += __shfl_xor(a0, off, 16);
    a1 += __shfl_xor(a1, off, 16);
    a2 += __shfl_xor(a2, off, 16);
    a3 += __shfl_xor(a3, off, 16);
    a4 += __shfl_xor(a4, off, 16);
  }
  if (p != 0) return;
  double* __restrict__ h = s.hess + 20 * (size_t)i + 5 * v;
  h[0] = a0;
  h[1] = a1;
  h[2] = a2;
  h[3] = a3;
  h[4] = a4;
  const double* Mi = m.moments + 9 * (size_t)i;
  s.pv[4 * (size_t)i + v] = qi - Mi[0] * a0 - Mi[1] * a1 - Mi[2] * a2 * 0.5 - Mi[3] * a3 * 0.5 - Mi[4] * a4;
}

// ---------------------------------------------------------------------------
// K4: Barth limiter, node-centric (limiter.C:39-497).  Min/max over the
// edge-connected ring, then the node's own unlimited reconstruction at each
// incident face point (k=1) or the 3 Gauss points (k=2 phase); one scalar
// per node (min over the 4 variables).
// ---------------------------------------------------------------------------
// One face-point candidate: r = (Qmax - q)/(qf - q) above, (Qmin - q)/(qf - q)
// below; Barth takes min(1, r) (limiter.C:154-246), Venkatakrishnan
// (r^2 + 2r)/(r^2 + r + 2) (limiter.C:696-730).  The numerator is selected
// first so a wave divides once per candidate whichever side its lanes take;
// the quotient is the same operation on the same operands as in the
// two-branch form.
//
// Barth's phi is min(1, min of the quotients), and IEEE division rounds
// monotonically, so a quotient can lower the running minimum ph (<= 1, a
// representable value) only if num/d < ph exactly.  fma(ph, d, -num) rounds
// ph d - num once, which keeps its sign; the division runs only when that
// sign allows a smaller quotient (or is zero, or the operands are not
// finite), so the minimum and every bit of phi are those of dividing every
// candidate, and a smooth region divides nowhere.
template <bool kVenkat>
__device__ __forceinline__ void limiter_candidate(double q, double qmin, double qmax, double qf, double* ph) {
  const double d = qf - q;
  const double num = (d > 0.) ? (qmax - q) : (qmin - q);
  if (!kVenkat) {
    if (!(d > 0. || d < 0.)) return;   // the candidate is 1: no effect on ph <= 1
    const double r = fma(*ph, d, -num);
    if ((d > 0. && r < 0.) || (d < 0. && r > 0.)) return;
    *ph = dmin(*ph, dmin(1.0, num / d));
    return;
  }
  double t = num / d;
  t = (t * t + 2. * t) / (t * t + t + 2.);
  *ph = dmin(*ph, (d > 0. || d < 0.) ? t : 1.0);
}

__device__ __forceinline__ void candidate(bool venkat, double q, double qmin, double qmax, double qf, double* ph) {
  if (venkat)
    limiter_candidate<true>(q, qmin, qmax, qf, ph);
  else
    limiter_candidate<false>(q, qmin, qmax, qf, ph);
}

// 16 lanes per node: lane v + 4p limits variable v over the node's edges and
// face points p, p+4, ...; min and max are exact, so combining the 4 parts
// and then the 4 variables (limiter.C:410-420) with shuffles gives the
// sequential loops' values.  An instance without a limiter this iteration
// gets phi = 1 (solve.C:343-347).  Mode bits (InstParams::lim): 1 limiter
// on, 2 Gauss points of the k=2 phase, 4 Venkatakrishnan's function instead
// of Barth's, 8 frozen (phi kept).
__global__ void __launch_bounds__(kMaxBlock) k_limiter(DevMesh m, DevState s) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  const bool own = m.lim_chunks != nullptr;
  int n = work_elem(own ? m.lim_chunks : m.node_chunks, own ? m.n_lim_chunks : m.n_node_chunks, tid >> 4, m.nn);
  const int v = tid & 3, p = (tid >> 2) & 3;
  if (n >= m.nn) return;
  const InstParams& ip = s.ip[inst_of(m, m.inst_node, n)];
  if (!ip.active) return;
  const int lim = ip.lim;
  if (lim & LIM_FROZEN) return;
  double* __restrict__ phi = s.phi;
  if (!(lim & LIM_ON)) {
    if ((tid & 15) == 0) phi[n] = 1.0;
    return;
  }
  const double* __restrict__ nQ = s.nQ;
  const double q = nQ[4 * (size_t)n + v];
  double qmin = q, qmax = q, ph = 1.0;
  const int eb = m.nedge_off[n], ee = m.nedge_off[n + 1];
  const int sb = m.nsub_off[n], se = m.nsub_off[n + 1];
  const bool gauss = lim & LIM_GAUSS;
  // k=2 phase: the node's derivatives, point value and the Gauss points of
  // its first face incidences are requested before the neighbour min/max,
  // so the two load chains overlap (up to three incidences per lane and pass)
  constexpr int kU = 3;
  double xg[kU][6];
  double h0 = 0., h1 = 0., h2 = 0., h3 = 0., h4 = 0., p0 = 0.;
  auto load_xg = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = k0 + 4 * u;
      const double* src = m.inc_xg + 6 * (size_t)(k < se ? k : k0);
#pragma unroll
      for (int c = 0; c < 6; ++c) xg[u][c] = src[c];
    }
  };
  if (gauss) {
    const double* h = s.hess + 20 * (size_t)n + 5 * v;
    h0 = h[0]; h1 = h[1]; h2 = h[2]; h3 = h[3]; h4 = h[4];
    p0 = s.pv[4 * (size_t)n + v];
    if (sb + p < se) load_xg(sb + p);
  }
  // two edges per lane and pass, both loads in flight together
  for (int k = eb + p; k < ee; k += 8) {
    const bool two = k + 4 < ee;
    int o0 = m.inc_other[k] & 0x7fffffff;
    int o1 = two ? (m.inc_other[k + 4] & 0x7fffffff) : o0;
    double q0 = nQ[4 * (size_t)o0 + v], q1 = nQ[4 * (size_t)o1 + v];
    qmin = dmin(dmin(qmin, q0), q1);
    qmax = dmax(dmax(qmax, q0), q1);
  }
#pragma unroll
  for (int off = 4; off <= 8; off <<= 1) {
    qmin = dmin(qmin, __shfl_xor(qmin, off, 16));
    qmax = dmax(qmax, __shfl_xor(qmax, off, 16));
  }
  const double xn0 = m.x[2 * (size_t)n], xn1 = m.x[2 * (size_t)n + 1];
  const bool venkat = lim & LIM_VENKAT;
  if (gauss) {
    for (int k0 = sb + p; k0 < se; k0 += 4 * kU) {
      if (k0 != sb + p) load_xg(k0);
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (k0 + 4 * u >= se) break;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          double dx0 = xg[u][2 * g] - xn0;
          double dx1 = xg[u][2 * g + 1] - xn1;
          double qf = p0;
          qf += (h0 * dx0 + h1 * dx1 + h2 * dx0 * dx0 * 0.5 + h3 * dx1 * dx1 * 0.5 + h4 * dx0 * dx1);
          candidate(venkat, q, qmin, qmax, qf, &ph);
        }
      }
    }
  } else {
    // face points of the linear reconstruction (the limiter runs only in
    // the gradient phase of k=1)
    const double g0 = s.grad[8 * (size_t)n + 2 * v];
    const double g1 = s.grad[8 * (size_t)n + 2 * v + 1];
    for (int k = sb + p; k < se; k += 4) {
      double qf = q;
      double dx0 = m.inc_fc[2 * (size_t)k] - xn0, dx1 = m.inc_fc[2 * (size_t)k + 1] - xn1;
      qf += (g0 * dx0 + g1 * dx1);
      candidate(venkat, q, qmin, qmax, qf, &ph);
    }
  }
#pragma unroll
  for (int off = 1; off <= 8; off <<= 1) ph = dmin(ph, __shfl_xor(ph, off, 16));
  if ((tid & 15) == 0) phi[n] = ph;
}

// ---------------------------------------------------------------------------
// K5+K7: edge-centric face fluxes and forward-difference Jacobian blocks
// (residual.C:112-333 with Reconstruct, reconstruct.C:181-205, and
// compute_numerical_jacobian_roe, jacobian.C:45-185).  2 lanes per edge:
// both evaluate the face flux of the edge's 1-2 subedges in ascending order;
// lane 0 stores +f at the L incidence and perturbs the L state (block (R,L)),
// lane 1 stores -f at the R incidence and perturbs the R state (block
// (L,R)).  The unperturbed side's primitive state and physical flux are
// computed once per subedge and reused by the four perturbed evaluations,
// which are the same operations on the same operands as four full
// Roe_flux_centered calls.  Each lane owns its block (no atomics); each entry
// is accumulated over the subedges in the reference's order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxBlock) k_edge_flux(DevMesh m, DevState st) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  int e = work_elem(m.edge_chunks, m.n_edge_chunks, tid >> 1, m.ne), side = tid & 1;
  if (e >= m.ne) return;
  const InstParams& ip = st.ip[inst_of(m, m.inst_edge, e)];
  if (!ip.active) return;
  const int recon = ip.recon_flux;
  const bool jupdate = ip.jac;
  const double gamma = ip.gamma;
  const double* __restrict__ nQ = st.nQ;
  const double* __restrict__ grad = st.grad;
  const double* __restrict__ phi = st.phi;
  double* __restrict__ inc_flux = st.inc_flux;
  const double eps = 1.0e-08;
  int L = m.edges[2 * e], R = m.edges[2 * e + 1];
  double qL0[4], qR0[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    qL0[j] = nQ[4 * (size_t)L + j];
    qR0[j] = nQ[4 * (size_t)R + j];
  }
  double blk[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) blk[k] = 0.0;
  for (int t = 0; t < 2; ++t) {
    int s = m.edge_sub[2 * e + t];
    if (s < 0) break;
    double nx = m.sub_n[3 * (size_t)s], ny = m.sub_n[3 * (size_t)s + 1], len = m.sub_n[3 * (size_t)s + 2];
    const size_t kL = (size_t)m.sub_inc[2 * (size_t)s], kR = (size_t)m.sub_inc[2 * (size_t)s + 1];
    double ql[4], qr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { ql[j] = qL0[j]; qr[j] = qR0[j]; }
    if (recon == 1) {
      double fc0 = m.inc_fc[2 * kL], fc1 = m.inc_fc[2 * kL + 1];
      double dxl0 = fc0 - m.x[2 * (size_t)L], dxl1 = fc1 - m.x[2 * (size_t)L + 1];
      double dxr0 = fc0 - m.x[2 * (size_t)R], dxr1 = fc1 - m.x[2 * (size_t)R + 1];
      double phiL = phi[L], phiR = phi[R];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ql[j] += (phiL * grad[8 * (size_t)L + 2 * j] * dxl0 + phiL * grad[8 * (size_t)L + 2 * j + 1] * dxl1);
        qr[j] += (phiR * grad[8 * (size_t)R + 2 * j] * dxr0 + phiR * grad[8 * (size_t)R + 2 * j + 1] * dxr1);
      }
    }
    const Prim PL = prim(ql, gamma), PR = prim(qr, gamma);
    double FL[4], FR[4], f[4];
    euler_flux_prim(PL, FL, nx, ny, len);
    euler_flux_prim(PR, FR, nx, ny, len);
    roe_flux_parts(nx, ny, len, gamma, ql, qr, PL, PR, FL, FR, f);
    // R[L] += f and R[R] -= f become additions of +f and -f (bit-identical)
    double4* dst = reinterpret_cast<double4*>(inc_flux + 4 * (side ? kR : kL));
    *dst = side ? make_double4(-f[0], -f[1], -f[2], -f[3]) : make_double4(f[0], f[1], f[2], f[3]);
    if (jupdate) {
      // the side this lane perturbs and the side it keeps
      const double* qs = side ? qr : ql;
      const Prim& Pk = side ? PL : PR;
      const double* Fk = side ? FL : FR;
      const double* qk = side ? ql : qr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double qp[4], fp[4], Fp[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) qp[j] = qs[j];
        qp[i] += eps;
        const Prim Pp = prim(qp, gamma);
        euler_flux_prim(Pp, Fp, nx, ny, len);
        // (qp, qr) for the L side, (ql, qp) for the R side
        roe_flux_parts(nx, ny, len, gamma, side ? qk : qp, side ? qp : qk, side ? Pk : Pp, side ? Pp : Pk,
                       side ? Fk : Fp, side ? Fp : Fk, fp);
        // block (R,L) -= dF/dqL, block (L,R) += dF/dqR; x - d is x + (-d)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double d = (fp[j] - f[j]) / eps;
          blk[j * 4 + i] += side ? d : -d;
        }
      }
    }
  }
  // the edge owns its two off-diagonal blocks: written every pass (zeros
  // when the instance skips the Jacobian, as build_matrix clears them)
  double2* dst = reinterpret_cast<double2*>(st.LHS + 16 * (size_t)m.edge_slot[2 * e + (side ? 0 : 1)]);
#pragma unroll
  for (int k = 0; k < 8; ++k) dst[k] = make_double2(blk[2 * k], blk[2 * k + 1]);
}

// ---------------------------------------------------------------------------
// K6: boundary edges: characteristic ghost update with memory (cvbc.C:37-481),
// first-order Roe flux against the ghost (residual.C:1157-1167) and the
// diagonal approximate Jacobian (jacobian.C:681-962).  Each bedge owns its
// ghost, so the sequential reference loop is order-independent.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_bedge(DevMesh m, DevState s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nbe) return;
  const int inst = inst_of(m, m.inst_bedge, i);
  const InstParams& p = s.ip[inst];
  if (!p.active) return;
  const int jupdate = p.jac;
  double* __restrict__ nQ = s.nQ;
  double* __restrict__ bflux = s.bflux;
  double* __restrict__ bjac = s.bjac;
  int real = m.bedge_real[i];
  size_t gh = (size_t)m.nn + i;
  double nx = m.bedge_n[3 * (size_t)i], ny = m.bedge_n[3 * (size_t)i + 1], len = m.bedge_n[3 * (size_t)i + 2];
  double qi[4], qg[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) { qi[j] = nQ[4 * (size_t)real + j]; qg[j] = nQ[4 * gh + j]; }
  if (!ghost_state(m.bedge_bc[i], nx, ny, p.gamma, p.uinf, p.vinf, p.pinf, qi, qg)) atomicOr(s.err + inst, 1);
#pragma unroll
  for (int j = 0; j < 4; ++j) nQ[4 * gh + j] = qg[j];
  double f[4];
  roe_flux(nx, ny, len, p.gamma, qi, qg, f);
#pragma unroll
  for (int j = 0; j < 4; ++j) bflux[4 * (size_t)i + j] = f[j];
  if (jupdate) {
    double d[16];
    boundary_jacobian(nx, ny, len, p.gamma, qi, qg, d);
#pragma unroll
    for (int k = 0; k < 16; ++k) bjac[16 * (size_t)i + k] = d[k];
  }
}

// ---------------------------------------------------------------------------
// Node assembly: ev_con/dt and the V/dt diagonal (init_memory.C:227-303,
// cv_calc.C:819-953), residual gather in ascending subedge then bedge order,
// Jacobian diagonal = V/dt I - sum of the transposed off-diagonal blocks of
// the row (each interior face adds dF/dq_L to diag L and subtracts it from
// block (R,L)) + boundary blocks, RHS = -R.
// ---------------------------------------------------------------------------
// 4 lanes per node: lane v owns residual component v and row v of the
// diagonal block, so the quad reads each 128-byte block contiguously.
__global__ void __launch_bounds__(kMaxBlock) k_node(DevMesh m, DevState s, int flags) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  int n = work_elem(m.node_chunks, m.n_node_chunks, tid >> 2, m.nn), v = tid & 3;
  if (n >= m.nn) return;
  const int inst = inst_of(m, m.inst_node, n);
  const InstParams& p = s.ip[inst];
  if (!p.active) return;
  const bool evcon = flags & NODE_EVCON, interior = flags & NODE_INTERIOR, boundary = flags & NODE_BOUNDARY;
  const bool jac = (flags & NODE_JAC) && p.jac, rhs = flags & NODE_RHS;
  double q[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) q[j] = s.nQ[4 * (size_t)n + j];
  double Rv = interior ? 0.0 : s.R[4 * (size_t)n + v];
  double ev = 0.0;
  if (evcon || interior) {
#pragma unroll 4
    for (int k = m.nsub_off[n]; k < m.nsub_off[n + 1]; ++k) {
      if (evcon) {
        const double* nr = m.inc_n + 4 * (size_t)k;
        ev += (max_ev(q, nr[0], nr[1], p.gamma) * nr[2]);
      }
      if (interior) Rv += s.inc_flux[4 * (size_t)k + v];
    }
  }
  if (evcon || boundary) {
    for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
      int b = m.nbe_list[k];
      if (evcon) {
        double nx = m.bedge_n[3 * (size_t)b], ny = m.bedge_n[3 * (size_t)b + 1], len = m.bedge_n[3 * (size_t)b + 2];
        ev += (max_ev(q, nx, ny, p.gamma) * len);
      }
      if (boundary) Rv += s.bflux[4 * (size_t)b + v];
    }
  }
  if (flags & NODE_EVONLY) {
    if (v == 0) {
      if (!isfinite(ev)) atomicOr(s.err + inst, 2);
      s.ev_con[n] = ev;
    }
    return;
  }
  double D[4];
  double* Dg = s.LHS + 16 * (size_t)m.iau[n] + 4 * v;
  if (evcon) {
    double V = m.cv_area[n];
    double dtn = p.cfl * (V / ev);
    if (v == 0) {
      if (!(dtn > 0.0) || !isfinite(dtn) || !isfinite(ev)) atomicOr(s.err + inst, 2);
      s.ev_con[n] = ev;
      s.dt[n] = dtn;
    }
    double vd = V / dtn;
#pragma unroll
    for (int c = 0; c < 4; ++c) D[c] = (c == v) ? vd : 0.0;
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) D[c] = Dg[c];
  }
  if (interior && jac) {
#pragma unroll 4
    for (int q = m.roff[n]; q < m.roff[n + 1]; ++q) {
      const double* B = s.LHS + 16 * (size_t)m.tslot[q] + 4 * v;
#pragma unroll
      for (int c = 0; c < 4; ++c) D[c] -= B[c];
    }
  }
  if (boundary && jac) {
    for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
      const double* B = s.bjac + 16 * (size_t)m.nbe_list[k] + 4 * v;
#pragma unroll
      for (int c = 0; c < 4; ++c) D[c] += B[c];
    }
  }
  if (evcon || jac) {
#pragma unroll
    for (int c = 0; c < 4; ++c) Dg[c] = D[c];
  }
  if (interior || boundary) {
    s.R[4 * (size_t)n + v] = Rv;
    if (rhs) s.RHS[4 * (size_t)n + v] = -(Rv);
  }
}

__global__ void __launch_bounds__(kBlock) k_zero_offdiag(DevMesh m, double* __restrict__ LHS) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)m.nnz * 16) return;
  LHS[i] = 0.0;
}

// ---------------------------------------------------------------------------
// K8: multicolor block Gauss-Seidel (linear_algebra.C:122-401).  The 4x4
// diagonal blocks are factored once per solve (the reference refactors the
// same block every sweep with the same deterministic LU, so the factors are
// identical); each color's rows are independent and run in parallel; sweeps
// alternate forward and backward over the colors.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxBlock) k_diag_lu(DevMesh m, const InstParams* __restrict__ ip,
                                                    const double* __restrict__ LHS, double* __restrict__ LU,
                                                    double* __restrict__ dQ) {
  int t = work_elem(m.pos_chunks, m.n_pos_chunks, xcd_block(m.xcd_run) * blockDim.x + threadIdx.x, m.nn);
  if (t >= m.nn) return;
  int n = m.color_rows[t];
  if (!ip[inst_of(m, m.pos_inst, t)].active) return;
  double A[16];
  const double* D = LHS + 16 * (size_t)m.iau[n];
#pragma unroll
  for (int k = 0; k < 16; ++k) A[k] = D[k];
  lu4(A);
#pragma unroll
  for (int k = 0; k < 16; ++k) LU[16 * (size_t)t + k] = A[k];
#pragma unroll
  for (int j = 0; j < 4; ++j) dQ[4 * (size_t)n + j] = 0.0;
}

// rows [row_begin,row_end) of the color-sorted order; 4 lanes per row, lane v
// forms component v of RHS - sum_offdiag A dQ in the row's block order (a
// wave's 16 rows read block e of their rows as one contiguous 2 KB run of the
// sliced layout), then every lane of the quad solves with the factored
// diagonal.  The first kPre column ids are loaded together (the column array
// is padded, so without a degree test), then those blocks and neighbour
// states, before any arithmetic: the products are the same and are summed
// in the same order as one block at a time.  kDpp: lane v loads component v
// of each neighbour's dQ and the quad exchanges them (the same values).
template <int kPre, bool kDpp>
__global__ void __launch_bounds__(kMaxBlock) k_gs_color(DevMesh m, const InstParams* __restrict__ ip,
                                                     const double* __restrict__ LHS, const double* __restrict__ LU,
                                                     const double* __restrict__ RHS, double* __restrict__ dQ,
                                                     const int* __restrict__ chunks, int nchunks, int row_begin,
                                                     int row_end) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  // with a work list of nchunks chunks: row (tid / 4) % 16 of the wave's chunk
  int t = row_begin + (tid >> 2);
  if (chunks) {
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    t = w < nchunks ? chunks[w] + ((tid >> 2) & (kGsChunk - 1)) : row_end;
  }
  const int v = tid & 3;
  if (t >= row_end) return;
  int j = m.color_rows[t];
  int inst = inst_of(m, m.pos_inst, t);
  int deg = m.pdeg[t];
  const size_t base =
      m.slice_w ? (size_t)16 * m.slice_w * (size_t)(t >> 4) + (t & 15) : (size_t)m.sbase[t >> 4] + (t & 15);
  double r = RHS[4 * (size_t)j + v];
  // kPre 0: every block in the loop below, one at a time
  constexpr int kN = kPre > 0 ? kPre : 1;
  int nb[kN];
#pragma unroll
  for (int i = 0; i < kPre; ++i) nb[i] = m.ja[base + 16 * i];
  if (!ip[inst].active) return;
  double2 A01[kN], A23[kN], q01[kN], q23[kN];
  double qv[kN];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    if (i < deg) {
      const double2* Ab = reinterpret_cast<const double2*>(LHS + 16 * (base + 16 * i) + 4 * v);
      A01[i] = Ab[0];
      A23[i] = Ab[1];
      if (kDpp) {
        qv[i] = dQ[4 * (size_t)nb[i] + v];
      } else {
        const double2* qb = reinterpret_cast<const double2*>(dQ + 4 * (size_t)nb[i]);
        q01[i] = qb[0];
        q23[i] = qb[1];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    if (i < deg) {
      double q0, q1, q2, q3;
      if (kDpp) {
        q0 = quad_bcast_d<0>(qv[i]);
        q1 = quad_bcast_d<1>(qv[i]);
        q2 = quad_bcast_d<2>(qv[i]);
        q3 = quad_bcast_d<3>(qv[i]);
      } else {
        q0 = q01[i].x; q1 = q01[i].y; q2 = q23[i].x; q3 = q23[i].y;
      }
      r -= (A01[i].x * q0 + A01[i].y * q1 + A23[i].x * q2 + A23[i].y * q3);
    }
  }
  for (int e = kPre; e < deg; ++e) {
    size_t k = base + 16 * (size_t)e;
    const double* A = LHS + 16 * k + 4 * v;
    const double* q = dQ + 4 * (size_t)m.ja[k];
    r -= (A[0] * q[0] + A[1] * q[1] + A[2] * q[2] + A[3] * q[3]);
  }
  double rr[4];
  rr[0] = __shfl(r, 0, 4);
  rr[1] = __shfl(r, 1, 4);
  rr[2] = __shfl(r, 2, 4);
  rr[3] = __shfl(r, 3, 4);
  double Ld[16], X[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) Ld[k] = LU[16 * (size_t)t + k];
  lu_solve4(Ld, rr, X);
  double xv = v == 0 ? X[0] : v == 1 ? X[1] : v == 2 ? X[2] : X[3];
  double old = dQ[4 * (size_t)j + v];
  double dx = xv - old;
  dQ[4 * (size_t)j + v] = old + dx;
}

// The same sweep with one thread per row (all four components): the same
// products summed in the same order as the 4-lane form.
__global__ void __launch_bounds__(kMaxBlock) k_gs_row(DevMesh m, const InstParams* __restrict__ ip,
                                                      const double* __restrict__ LHS, const double* __restrict__ LU,
                                                      const double* __restrict__ RHS, double* __restrict__ dQ,
                                                      const int* __restrict__ chunks, int nchunks, int row_begin,
                                                      int row_end) {
  const int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  // with a work list of nchunks chunks: row tid % 16 of chunk tid / 16
  int t = row_begin + tid;
  if (chunks) t = tid / kGsChunk < nchunks ? chunks[tid / kGsChunk] + (tid & (kGsChunk - 1)) : row_end;
  if (t >= row_end) return;
  int j = m.color_rows[t];
  int inst = inst_of(m, m.pos_inst, t);
  int deg = m.pdeg[t];
  size_t base = (size_t)m.sbase[t >> 4] + (t & 15);
  double r[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) r[v] = RHS[4 * (size_t)j + v];
  if (!ip[inst].active) return;
  for (int e = 0; e < deg; ++e) {
    size_t k = base + 16 * (size_t)e;
    const double* A = LHS + 16 * k;
    const double* q = dQ + 4 * (size_t)m.ja[k];
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
#pragma unroll
    for (int v = 0; v < 4; ++v) r[v] -= (A[4 * v] * q0 + A[4 * v + 1] * q1 + A[4 * v + 2] * q2 + A[4 * v + 3] * q3);
  }
  double Ld[16], X[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) Ld[k] = LU[16 * (size_t)t + k];
  lu_solve4(Ld, r, X);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double old = dQ[4 * (size_t)j + v];
    double dx = X[v] - old;
    dQ[4 * (size_t)j + v] = old + dx;
  }
}

// nQ += dQ over the real nodes (solve.C:901-907)
__global__ void __launch_bounds__(kBlock) k_update(DevMesh m, const InstParams* __restrict__ ip,
                                                   double* __restrict__ nQ, const double* __restrict__ dQ) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = work_elem(m.node_chunks, m.n_node_chunks, k >> 2, m.nn);
  if (n >= m.nn) return;
  if (!ip[inst_of(m, m.inst_node, n)].active) return;
  const size_t i = 4 * (size_t)n + (k & 3);
  nQ[i] += dQ[i];
}

__global__ void __launch_bounds__(kBlock) k_copy(size_t n, double* __restrict__ dst, const double* __restrict__ src) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// the listed nodes' 4 words (one thread per word)
__global__ void __launch_bounds__(kBlock) k_copy_nodes(DevMesh m, double* __restrict__ dst,
                                                       const double* __restrict__ src) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = work_elem(m.node_chunks, m.n_node_chunks, k >> 2, m.nn);
  if (n >= m.nn) return;
  const size_t i = 4 * (size_t)n + (k & 3);
  dst[i] = src[i];
}

// k_copy_nodes which also keeps the elements of instances whose k=2 fields
// are deferred this step in `save`
__global__ void __launch_bounds__(kBlock) k_copy_nodes_save(DevMesh m, const InstParams* __restrict__ ip,
                                                            double* __restrict__ dst, const double* __restrict__ src,
                                                            double* __restrict__ save) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = work_elem(m.node_chunks, m.n_node_chunks, k >> 2, m.nn);
  if (n >= m.nn) return;
  const size_t i = 4 * (size_t)n + (k & 3);
  const double v = src[i];
  dst[i] = v;
  const InstParams& p = ip[inst_of(m, m.inst_node, n)];
  if (p.active && p.k2_deferred) save[i] = v;
}
// the ghosts: ghost g is the outer state of boundary edge g
__global__ void __launch_bounds__(kBlock) k_copy_ghosts_save(DevMesh m, const InstParams* __restrict__ ip,
                                                             double* __restrict__ dst, const double* __restrict__ src,
                                                             double* __restrict__ save) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 4 * m.ng) return;
  const int g = k >> 2;
  const size_t i = 4 * ((size_t)m.nn + g) + (k & 3);
  const double v = src[i];
  dst[i] = v;
  if (g >= m.nbe) return;
  const InstParams& p = ip[inst_of(m, m.inst_bedge, g)];
  if (p.active && p.k2_deferred) save[i] = v;
}

// one thread per real node or ghost; its instance's mode bits pick the copies
__global__ void __launch_bounds__(kBlock) k_checkpoint(DevMesh m, double* __restrict__ Q, double* __restrict__ Qsave,
                                                       double* __restrict__ Qphase, const int* __restrict__ mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.nn + m.nbe) return;
  const int md = mode[t < m.nn ? inst_of(m, m.inst_node, t) : inst_of(m, m.inst_bedge, t - m.nn)];
  if (md == 0) return;
  const size_t o = 4 * (size_t)t;
  if (md & (CKPT_RESTORE | CKPT_RESTORE_PHASE)) {
    const double* __restrict__ src = ((md & CKPT_RESTORE) ? Qsave : Qphase) + o;
#pragma unroll
    for (int k = 0; k < 4; ++k) Q[o + k] = src[k];
    return;
  }
  double q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = Q[o + k];
  if (md & CKPT_SAVE)
#pragma unroll
    for (int k = 0; k < 4; ++k) Qsave[o + k] = q[k];
  if (md & CKPT_SAVE_PHASE)
#pragma unroll
    for (int k = 0; k < 4; ++k) Qphase[o + k] = q[k];
}

// Deterministic per-instance sums of R^2 and dQ^2: one block per chunk of
// kNormChunkNodes nodes (fixed tree order), then one wave per instance adds
// its chunks, lane l taking chunks l, l+64, ... counted from the instance's
// first chunk, followed by a fixed butterfly.  Both orders depend only on the
// instance's own node count, so an instance's norms are the same alone and
// in a batch.
__global__ void __launch_bounds__(kBlock) k_norm_chunk(NormChunks c, const InstParams* __restrict__ ip,
                                                       const double* __restrict__ R, const double* __restrict__ dQ,
                                                       double* __restrict__ part) {
  __shared__ double sr[kBlock], sd[kBlock];
  int ch = blockIdx.x;
  if (!ip[c.chunk_inst[ch]].active) return;
  size_t b = 4 * (size_t)c.chunk_begin[ch], e = 4 * (size_t)c.chunk_end[ch];
  double a = 0.0, d = 0.0;
  for (size_t i = b + threadIdx.x; i < e; i += kBlock) {
    a += R[i] * R[i];
    d += dQ[i] * dQ[i];
  }
  sr[threadIdx.x] = a;
  sd[threadIdx.x] = d;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (size_t)ch] = sr[0];
    part[2 * (size_t)ch + 1] = sd[0];
  }
}

__global__ void __launch_bounds__(kBlock) k_norm_inst(int ninst, NormChunks c, const InstParams* __restrict__ ip,
                                                      const double* __restrict__ part, double* __restrict__ out) {
  int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, l = threadIdx.x & 63;
  if (i >= ninst || !ip[i].active) return;
  double a = 0.0, d = 0.0;
  for (int ch = c.inst_chunk0[i] + l; ch < c.inst_chunk0[i + 1]; ch += 64) {
    a += part[2 * (size_t)ch];
    d += part[2 * (size_t)ch + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    d += __shfl_xor(d, off, 64);
  }
  if (l == 0) {
    out[2 * (size_t)i] = a;
    out[2 * (size_t)i + 1] = d;
  }
}

// ---------------------------------------------------------------------------
// Body forces (grid_io.C:1390-1697, straight boundaries): one thread per
// boundary edge; the three Gauss points of the bedge's half segment, the
// boundary reconstruction of Reconstruct_Gauss_Boundary (reconstruct.C:
// 1708-1890, PV_EXT) and the pressure, weighted 5/9, 8/9, 5/9 and scaled by
// len/2.  The per-bedge terms are summed in bedge order on the host, which
// keeps the reference's sequential sum.  Bedges that are not walls are
// skipped there.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_body_terms(DevMesh m, DevState s, const ForceParams* __restrict__ fp) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m.nbe) return;
  const int bc = m.bedge_bc[b];
  if (bc == 0 || (bc % 2 == 0)) {
    s.bforce[2 * (size_t)b] = 0.0;
    s.bforce[2 * (size_t)b + 1] = 0.0;
    return;
  }
  const ForceParams f = fp[inst_of(m, m.inst_bedge, b)];
  const int real = m.bedge_real[b];
  // straight wall: the bedge normal at every Gauss point, points stored
  // (+sqrt(3/5), 0, -sqrt(3/5)); curved wall (grid_io.C:1612-1660): the
  // curve's normal at each point, points stored (0, +sqrt(3/5), -sqrt(3/5))
  const bool curved = bc > 10;
  const double len = m.bedge_n[3 * (size_t)b + 2];
  const double x0 = m.x[2 * (size_t)real], x1 = m.x[2 * (size_t)real + 1];
  const double phi = s.phi[real], sigma = phi;
  double base[4], c1[4][2], c2[4][3];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    base[j] = (f.rorder == 3) ? s.pv[4 * (size_t)real + j] : s.nQ[4 * (size_t)real + j];
    if (f.rorder == 1) {
      c1[j][0] = s.grad[8 * (size_t)real + 2 * j];
      c1[j][1] = s.grad[8 * (size_t)real + 2 * j + 1];
    } else if (f.rorder == 3) {
      const double* h = s.hess + 20 * (size_t)real + 5 * j;
      c1[j][0] = h[0];
      c1[j][1] = h[1];
      c2[j][0] = h[2];
      c2[j][1] = h[3];
      c2[j][2] = h[4];
    }
  }
  const double w[3] = {curved ? 8. / 9. : 5. / 9., curved ? 5. / 9. : 8. / 9., 5. / 9.};
  double temp1 = 0., temp2 = 0.;
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const double nx = -m.bedge_gn[6 * (size_t)b + 2 * g], ny = -m.bedge_gn[6 * (size_t)b + 2 * g + 1];
    const double dx0 = m.bedge_xg[6 * (size_t)b + 2 * g] - x0;
    const double dx1 = m.bedge_xg[6 * (size_t)b + 2 * g + 1] - x1;
    double q[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      q[j] = base[j];
      if (f.rorder == 1)
        q[j] += (phi * c1[j][0] * dx0 + phi * c1[j][1] * dx1);
      else if (f.rorder == 3)
        q[j] += (phi * c1[j][0] * dx0 + phi * c1[j][1] * dx1 + sigma * c2[j][0] * dx0 * dx0 * 0.5 +
                 sigma * c2[j][1] * dx1 * dx1 * 0.5 + sigma * c2[j][2] * dx0 * dx1);
    }
    const double r = q[0], u = q[1] / r, v = q[2] / r, e = q[3];
    const double P = (f.gamma - 1.) * (e - 0.5 * r * (u * u + v * v));
    temp1 += (w[g] * (P * ny * f.ca - P * nx * f.sa));
    temp2 += (w[g] * (P * nx * f.ca + P * ny * f.sa));
  }
  s.bforce[2 * (size_t)b] = temp1 * len * 0.5;
  s.bforce[2 * (size_t)b + 1] = temp2 * len * 0.5;
}

// Streaming probes of the device's achievable HBM bandwidth (no counterpart
// in the reference): a read of n float4 with 8 loads in flight per thread,
// and a copy.
__global__ void __launch_bounds__(kBlock) k_probe_read(const float4* __restrict__ a, size_t n, float* __restrict__ out) {
  constexpr int U = 8;
  size_t i = (size_t)blockIdx.x * blockDim.x * U + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x * U;
  float acc = 0.f;
  for (; i + (U - 1) * (size_t)blockDim.x < n; i += stride) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[i + u * (size_t)blockDim.x];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  // never true for the zero-filled probe buffer; keeps the loads live
  if (acc == 1.0f) out[0] = acc;
}
__global__ void __launch_bounds__(kBlock) k_probe_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}
void launch_probe_read(const void* a, size_t n16, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_read, dim3(8192), dim3(kBlock), 0, st, (const float4*)a, n16, out);
}
void launch_probe_copy(const void* a, void* b, size_t n16, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_copy, dim3(8192), dim3(kBlock), 0, st, (const float4*)a, (float4*)b, n16);
}

template <int kForm>
__global__ void __launch_bounds__(kBlock) k_roe_batch(int n, double gamma, const double* __restrict__ nrm,
                                                      const double* __restrict__ qL, const double* __restrict__ qR,
                                                      double* __restrict__ flux) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[4], b[4], f[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) { a[j] = qL[4 * (size_t)i + j]; b[j] = qR[4 * (size_t)i + j]; }
  if (kForm == 0)
    roe_flux(nrm[3 * (size_t)i], nrm[3 * (size_t)i + 1], nrm[3 * (size_t)i + 2], gamma, a, b, f);
  else
    roe_flux_wave(nrm[3 * (size_t)i], nrm[3 * (size_t)i + 1], nrm[3 * (size_t)i + 2], gamma, a, b, f);
#pragma unroll
  for (int j = 0; j < 4; ++j) flux[4 * (size_t)i + j] = f[j];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_copy(double* dst, const double* src, size_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, dst, src);
}
void launch_copy_state(const DevMesh& m, double* dst, const double* src, hipStream_t st) {
  if (!m.node_chunks) {
    launch_copy(dst, src, ((size_t)m.nn + m.ng) * 4, st);
    return;
  }
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn > 0) hipLaunchKernelGGL(k_copy_nodes, dim3(nblocks(4 * nn)), dim3(kBlock), 0, st, m, dst, src);
  launch_copy(dst + 4 * (size_t)m.nn, src + 4 * (size_t)m.nn, 4 * (size_t)m.ng, st);
}
void launch_copy_state_save(const DevMesh& m, const InstParams* ip, double* dst, const double* src, double* save,
                            hipStream_t st) {
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn > 0) hipLaunchKernelGGL(k_copy_nodes_save, dim3(nblocks(4 * nn)), dim3(kBlock), 0, st, m, ip, dst, src, save);
  if (m.ng > 0)
    hipLaunchKernelGGL(k_copy_ghosts_save, dim3(nblocks(4 * m.ng)), dim3(kBlock), 0, st, m, ip, dst, src, save);
}
void launch_checkpoint(const DevMesh& m, double* Q, double* Qsave, double* Qphase, const int* mode, hipStream_t st) {
  if (m.nn + m.nbe == 0) return;
  hipLaunchKernelGGL(k_checkpoint, dim3(nblocks(m.nn + m.nbe)), dim3(kBlock), 0, st, m, Q, Qsave, Qphase, mode);
}
void launch_gradient(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  hipLaunchKernelGGL(k_gradient, dim3(nblocks(nn, m.tune.cell_block)), dim3(m.tune.cell_block), 0, st, m, s);
}
void launch_derivatives(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  if (m.tune.deriv_lds && m.ndtile > 0)
    hipLaunchKernelGGL(k_derivatives_lds, dim3(m.tile_list ? m.n_tile_list : m.ndtile), dim3(16 * kDerivTileNodes),
                       0, st, m, s);
  else
    hipLaunchKernelGGL(k_derivatives,
                       dim3(nblocks(16 * work_count(m.node_chunks, m.n_node_chunks, m.nn), m.tune.cell_block)),
                       dim3(m.tune.cell_block), 0, st, m,
                       s);
}
void launch_limiter(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = m.lim_chunks ? kWorkChunk * m.n_lim_chunks : work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  hipLaunchKernelGGL(k_limiter, dim3(nblocks(16 * nn, m.tune.cell_block)), dim3(m.tune.cell_block), 0, st, m, s);
}
void launch_edge_flux(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.ne == 0) return;
  const int ne = work_count(m.edge_chunks, m.n_edge_chunks, m.ne);
  if (ne == 0) return;
  hipLaunchKernelGGL(k_edge_flux, dim3(nblocks(2 * ne, m.tune.face_block)), dim3(m.tune.face_block), 0, st, m, s);
}
void launch_bedge(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nbe == 0) return;
  hipLaunchKernelGGL(k_bedge, dim3(nblocks(m.nbe)), dim3(kBlock), 0, st, m, s);
}
void launch_node(const DevMesh& m, const DevState& s, int flags, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  hipLaunchKernelGGL(k_node, dim3(nblocks(4 * nn, m.tune.node_block)), dim3(m.tune.node_block), 0, st, m, s, flags);
}
void launch_zero_offdiag(const DevMesh& m, const DevState& s, hipStream_t st) {
  size_t n = (size_t)m.nnz * 16;
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_offdiag, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, m, s.LHS);
}
// Color visiting order of `subits` symmetric sweeps (even sweeps ascend, odd
// sweeps descend).  A color repeated back to back at a turnaround is dropped
// (the reference's row-sequential sweep has no such repeat); the backward
// flag lets instances with fewer colors drop their own turnaround repeat.
int gs_schedule(int ncolor, int subits, GsLaunch* out) {
  int n = 0, last = -1;
  for (int it = 0; it < subits; ++it)
    for (int cc = 0; cc < ncolor; ++cc) {
      int c = (it % 2 == 0) ? cc : ncolor - 1 - cc;
      if (c == last) continue;
      out[n].color = c;
      out[n].backward = it % 2;
      ++n;
      last = c;
    }
  return n;
}

void launch_diag_lu(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int np = work_count(m.pos_chunks, m.n_pos_chunks, m.nn);
  if (np == 0) return;
  hipLaunchKernelGGL(k_diag_lu, dim3(nblocks(np, m.tune.cell_block)), dim3(m.tune.cell_block), 0, st, m, s.ip, s.LHS,
                     s.LU, s.dQ);
}

namespace {
template <int kPre>
void gs_launch(const DevMesh& m, const DevState& s, dim3 grid, dim3 block, const int* ch, int nch, int b, int e,
               hipStream_t st) {
  if (kPre == 0) {   // no prefetch: the gather form does not apply
    hipLaunchKernelGGL((k_gs_color<0, false>), grid, block, 0, st, m, s.ip, s.LHS, s.LU, s.RHS, s.dQ, ch, nch, b, e);
    return;
  }
  if (m.tune.gs32_dpp)
    hipLaunchKernelGGL((k_gs_color<kPre, true>), grid, block, 0, st, m, s.ip, s.LHS, s.LU, s.RHS, s.dQ, ch, nch, b, e);
  else
    hipLaunchKernelGGL((k_gs_color<kPre, false>), grid, block, 0, st, m, s.ip, s.LHS, s.LU, s.RHS, s.dQ, ch, nch, b, e);
}
}  // namespace

void launch_gs_sweeps(const DevMesh& m, const DevState& s, const ColorRanges& cr, int subits, hipStream_t st) {
  std::vector<GsLaunch> sched((size_t)subits * m.ncolor + 1);
  int ns = gs_schedule(m.ncolor, subits, sched.data());
  for (int i = 0; i < ns; ++i) {
    int c = sched[i].color;
    int b = cr.off[c], e = sched[i].backward ? b + cr.nback[c] : cr.off[c + 1];
    if (e <= b) continue;
    const int bs = m.tune.gs_block;
    const int* ch = nullptr;
    int nch = 0, nrows = e - b;
    if (cr.chunks) {
      const int r = 2 * c + sched[i].backward;
      ch = cr.chunks + cr.chunk_off[r];
      nch = cr.chunk_off[r + 1] - cr.chunk_off[r];
      nrows = kGsChunk * nch;
      if (nrows == 0) continue;
    }
    if (m.tune.gs_lanes == 1)
      hipLaunchKernelGGL(k_gs_row, dim3(nblocks(nrows, bs)), dim3(bs), 0, st, m, s.ip, s.LHS, s.LU, s.RHS, s.dQ, ch,
                         nch, b, e);
    else if (m.tune.gs32_pre == 0)
      gs_launch<0>(m, s, dim3(nblocks(4 * nrows, bs)), dim3(bs), ch, nch, b, e, st);
    else if (m.tune.gs32_pre <= 4)
      gs_launch<4>(m, s, dim3(nblocks(4 * nrows, bs)), dim3(bs), ch, nch, b, e, st);
    else if (m.tune.gs32_pre <= 6)
      gs_launch<6>(m, s, dim3(nblocks(4 * nrows, bs)), dim3(bs), ch, nch, b, e, st);
    else
      gs_launch<8>(m, s, dim3(nblocks(4 * nrows, bs)), dim3(bs), ch, nch, b, e, st);
  }
}
void launch_update(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  hipLaunchKernelGGL(k_update, dim3(nblocks(4 * nn)), dim3(kBlock), 0, st, m, s.ip, s.nQ, s.dQ);
}
void launch_norms(const DevMesh& m, const DevState& s, const NormChunks& c, hipStream_t st) {
  if (c.nchunk > 0)
    hipLaunchKernelGGL(k_norm_chunk, dim3(c.nchunk), dim3(kBlock), 0, st, c, s.ip, s.R, s.dQ, s.chunk_part);
  hipLaunchKernelGGL(k_norm_inst, dim3(nblocks(64 * m.ninst)), dim3(kBlock), 0, st, m.ninst, c, s.ip, s.chunk_part,
                     s.inst_norms);
}
void launch_body_terms(const DevMesh& m, const DevState& s, const ForceParams* fp, hipStream_t st) {
  if (m.nbe == 0) return;
  hipLaunchKernelGGL(k_body_terms, dim3(nblocks(m.nbe)), dim3(kBlock), 0, st, m, s, fp);
}
void launch_roe_batch(int n, int form, double gamma, const double* normals, const double* qL, const double* qR,
                      double* flux, hipStream_t st) {
  if (n <= 0) return;
  if (form == 0)
    hipLaunchKernelGGL(k_roe_batch<0>, dim3(nblocks(n)), dim3(kBlock), 0, st, n, gamma, normals, qL, qR, flux);
  else
    hipLaunchKernelGGL(k_roe_batch<1>, dim3(nblocks(n)), dim3(kBlock), 0, st, n, gamma, normals, qL, qR, flux);
}

}  // namespace fsk
