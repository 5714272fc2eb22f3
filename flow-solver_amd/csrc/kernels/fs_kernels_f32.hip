// Kernels of the mixed-precision path (gfx950, wave64).
//
// The outer loop stays FP64: the node pass k_resid_jac32 evaluates the FP64
// residual (Roe flux with the dissipation in closed form, at the FP64 face
// states), ev_con and dt; the limiter, gradient, derivatives, norms,
// convergence test and state update are the FP64 kernels.  Inside it the
// Jacobian blocks and the Gauss-Seidel sweeps run in FP32: the analytic Roe
// Jacobian of compute_approximate_jacobian_roe (jacobian.C:301-658) at the
// same reconstructed face states the reference differentiates around, built
// in the same node pass from the flux's Roe quantities, FP32 block storage
// (64 B per 4x4 block), diagonal blocks inverted in FP64 with pivoting and
// stored FP32, and FP32 sweeps.  The fixed point is the FP64 residual's, so
// converged states agree with the FP64 path to the FP64 stopping tolerance.
//
// Contraction (FMA) is on in this file, for the mixed path's FP64 residual
// too: unlike the FP64 path's kernels it is not bound to the reference's
// operation order, and contraction moves its fluxes by last-bit amounts only.
#include "fs_device.hpp"
#include "fs_kernels.hpp"

#include <vector>

#pragma clang fp contract(fast)

namespace fsk {
using namespace fsd;

namespace {
constexpr int kBlock = 256;
inline int nblocks(int n, int b = kBlock) { return (n + b - 1) / b; }
}  // namespace

// Exchange within a lane quad through DPP quad permutes (no LDS round trip):
// xor 1 swaps lanes (0,1) and (2,3), xor 2 swaps (0,2) and (1,3).  Every lane
// of the quad must be active.
template <int kXor>
__device__ __forceinline__ int quad_xor_i(int x) {
  constexpr int ctrl = kXor == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
  return __builtin_amdgcn_update_dpp(0, x, ctrl, 0xF, 0xF, false);
}
template <int kXor>
__device__ __forceinline__ float quad_xor(float x) {
  return __int_as_float(quad_xor_i<kXor>(__float_as_int(x)));
}
template <int kXor>
__device__ __forceinline__ double quad_xor(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = quad_xor_i<kXor>((int)(b & 0xffffffffLL)), hi = quad_xor_i<kXor>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Lane r's value in every lane of its quad (DPP quad_perm [r,r,r,r]).
template <int kLane>
__device__ __forceinline__ float quad_bcast(float x) {
  constexpr int ctrl = kLane * 0x55;
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, 0xF, 0xF, false));
}

// Euler flux Jacobian dF.n/dq (jacobian.C:561-582) in FP32 from the state's
// velocity and total energy per mass (e = E/rho)
__device__ __forceinline__ void flux_jacobian_f32(float nx, float ny, float g, float u, float v, float e, float* d) {
  const float th = u * nx + v * ny;
  const float ph = 0.5f * (g - 1.0f) * (u * u + v * v);
  const float gE = g * e;
  d[0] = 0.0f; d[1] = nx; d[2] = ny; d[3] = 0.0f;
  d[4] = ph * nx - u * th; d[5] = (2.0f - g) * u * nx + th; d[6] = u * ny - (g - 1.0f) * v * nx; d[7] = (g - 1.0f) * nx;
  d[8] = ph * ny - v * th; d[9] = v * nx - (g - 1.0f) * u * ny; d[10] = th + (2.0f - g) * v * ny; d[11] = (g - 1.0f) * ny;
  d[12] = (2.0f * ph - gE) * th; d[13] = (1.0f - g) * u * th + nx * (gE - ph);
  d[14] = (1.0f - g) * v * th + ny * (gE - ph); d[15] = g * th;
}

// Inverse of a 4x4 block in FP64 by Gauss-Jordan elimination with partial
// pivoting, stored in FP32.  A factorization without pivoting rounded to FP32
// loses the block where its leading pivot is small against the other entries
// (the continuity row's diagonal is V/dt plus dissipation only, since the
// flux Jacobian's (0,0) entry is zero); the pivoted FP64 inverse is accurate
// to FP32 rounding of its entries.  Row selection uses compile-time indices
// only, so the augmented matrix stays in registers.
__device__ __forceinline__ void inv4_pivoted(const float* A, double* out) {
  double a[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[i][j] = (double)A[4 * i + j];
      a[i][4 + j] = (i == j) ? 1.0 : 0.0;
    }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double best = fabs(a[k][k]);
    int p = k;
#pragma unroll
    for (int r = k + 1; r < 4; ++r)
      if (fabs(a[r][k]) > best) { best = fabs(a[r][k]); p = r; }
#pragma unroll
    for (int r = k + 1; r < 4; ++r)
      if (r == p) {
#pragma unroll
        for (int c = 0; c < 8; ++c) { double tmp = a[k][c]; a[k][c] = a[r][c]; a[r][c] = tmp; }
      }
    const double ip = 1.0 / a[k][k];
#pragma unroll
    for (int c = 0; c < 8; ++c) a[k][c] *= ip;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (r == k) continue;
      const double f = a[r][k];
#pragma unroll
      for (int c = 0; c < 8; ++c) a[r][c] -= f * a[k][c];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[4 * i + j] = a[i][4 + j];
}

// Row r of the pre-multiplied sweep operands (r = 0..3, Di the pivoted FP64
// inverse): row r of Di B for an off-diagonal block B, and entry r of Di b
// for the right-hand side, FP64 products rounded once to FP32.  The diagonal
// pass and the fused node pass both call these, so the two forms of the
// operands agree bit for bit.
__device__ __forceinline__ float4 scale_block_row(const double* dr, const float* b) {
  double c0 = 0., c1 = 0., c2 = 0., c3 = 0.;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double d = dr[i];
    c0 += d * b[4 * i];
    c1 += d * b[4 * i + 1];
    c2 += d * b[4 * i + 2];
    c3 += d * b[4 * i + 3];
  }
  return make_float4((float)c0, (float)c1, (float)c2, (float)c3);
}
__device__ __forceinline__ float scale_rhs_entry(const double* dr, float4 rb) {
  return (float)(dr[0] * rb.x + dr[1] * rb.y + dr[2] * rb.z + dr[3] * rb.w);
}

// FP16 bit pattern of an FP32 value (round to nearest even); `ok` turns
// false when it does not fit the FP16 range
__device__ __forceinline__ unsigned int half_bits(float x, bool& ok) {
  ok = ok && fabsf(x) < 65520.0f;
  return (unsigned int)__builtin_bit_cast(unsigned short, (_Float16)x);
}

// per-node LDS stride (floats) of the fused node pass: the row's blocks and
// 4 floats of padding, so consecutive nodes start in different banks
__host__ __device__ __forceinline__ int fused_row_stride(int maxdeg) { return 16 * maxdeg + 4; }

// Residual and Jacobian pass of the mixed step, one pass per node (4 lanes
// per node), nothing re-read from an earlier pass but the boundary edges'
// fluxes and blocks:
//   * FP64 residual R = sum over the node's faces of the flux out of its
//     control volume (residual.C:112-333 scatter, gathered): each end node
//     evaluates the subedge's flux with its own state on the left and its
//     outward normal, which is the reference's f and -f up to rounding (the
//     two ends' fluxes cancel to a few ulps, not bit for bit); the sum order
//     is a fixed function of the node's own edge list (alone and in a batch
//     alike);
//   * ev_con (cv_calc.C:790-965) and dt = cfl V / ev_con (init_memory.C:227);
//   * FP32 analytic Jacobian (compute_approximate_jacobian_roe,
//     jacobian.C:301-658) at the same face states: with the outward normal
//     n_o, A = dF.n_o/dq and M = |A_roe| (built in FP32 from the FP64 flux's
//     Roe quantities; |A_roe| does not change with the normal's sign),
//       diagonal block of n      += len/2 (A(q_n) + M)
//       block (n, other)         += len/2 (A(q_other) - M)
//     so the row's off-diagonal blocks are written here once and the diagonal
//     needs no second read of them;  V/dt I + boundary blocks on the diagonal.
// Lane v of the node's quad handles subedge (v & 1) of edge incidences
// v >> 1, 2 + (v >> 1), ...; the two subedges of an edge meet in lanes v and
// v ^ 1, which add their block contributions and write the edge's block.
//
// kOut selects what the pass leaves for the sweeps:
//   0  the assembled FP32 blocks (off-diagonal in LHS32, diagonal in D32) and
//      right-hand side; the diagonal pass (k_diag_lu32) follows;
//   1  the pre-multiplied FP32 operands directly: the row's off-diagonal
//      blocks are staged in LDS instead of written, and once the diagonal is
//      complete every lane of the quad inverts it (inv4_pivoted, FP64), lane
//      v writes row v of Di B for each staged block B and entry v of Di b,
//      and zeroes dQ -- the same operations as k_diag_lu32<true>, so the
//      operands equal the two-pass form's bit for bit, without the second
//      pass's read and write of every block;
//   2  as 1, with the off-diagonal operands stored in FP16 (LHS16); a value
//      outside the FP16 range flags the instance (device error bit 4).
// The FP32 diagonal block is stored in every form.
template <int kOut>
__global__ void __launch_bounds__(kMaxBlock) k_resid_jac32(DevMesh m, DevState s) {
  extern __shared__ float s_rows[];
  const int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  const int n = work_elem(m.node_chunks, m.n_node_chunks, tid >> 2, m.nn), v = tid & 3;
  if (n >= m.nn) return;
  const int inst = inst_of(m, m.inst_node, n);
  const InstParams& p = s.ip[inst];
  if (!p.active) return;
  const int t = m.iau[n] - (m.nnz - m.nn);   // the row's color position
  const int jac = p.jac, recon = p.recon_flux;
  const double g = p.gamma, g1 = p.gamma - 1.0;
  const float gf = (float)p.gamma, gm1 = (float)(p.gamma - 1.0);
  const double* __restrict__ nQ = s.nQ;
  const double* __restrict__ grad = s.grad;
  PrimW Pn;
  {
    double qn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) qn[j] = nQ[4 * (size_t)n + j];
    Pn = prim_w(qn, g1);
  }
  const double cn = sqrt((g * Pn.p) * Pn.ir);   // the node's own sound speed (cv_calc.C:819-849)
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, ev = 0.0;
  float dg[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) dg[c] = 0.0f;
  const int sub = v & 1;
  const int k0 = m.nedge_off[n], k1 = m.nedge_off[n + 1];
  // the next face's index, geometry and neighbour state are loaded one
  // iteration ahead, so their latency overlaps the current face's arithmetic
  int k = k0 + (v >> 1);
  int ent_n = 0;
  double geo_n[3] = {0.0, 0.0, 0.0}, qb_n[4] = {1.0, 0.0, 0.0, 1.0};
  // linear reconstruction (recon 1): the node's own gradient, limiter value
  // and coordinates once, and the next face's face point and the other
  // node's gradient, limiter value and coordinates one iteration ahead
  double ga[8], xa[2], phia = 0.0;
  double gb_n[8], xb_n[2] = {0.0, 0.0}, fc_n[2] = {0.0, 0.0}, phib_n = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) gb_n[j] = 0.0;
  auto load_face = [&](int kk) {
    ent_n = m.inc_other[kk];
    const double* geo = m.einc_geo + 6 * (size_t)kk + 3 * sub;
#pragma unroll
    for (int c = 0; c < 3; ++c) geo_n[c] = geo[c];
    const size_t on = (size_t)(ent_n & 0x7fffffff);
#pragma unroll
    for (int j = 0; j < 4; ++j) qb_n[j] = nQ[4 * on + j];
    if (recon == 1) {
      const double* fc = m.einc_fc + 4 * (size_t)kk + 2 * sub;
      fc_n[0] = fc[0];
      fc_n[1] = fc[1];
      phib_n = s.phi[on];
      xb_n[0] = m.x[2 * on];
      xb_n[1] = m.x[2 * on + 1];
#pragma unroll
      for (int j = 0; j < 8; ++j) gb_n[j] = grad[8 * on + j];
    }
  };
  if (recon == 1) {
    phia = s.phi[n];
    xa[0] = m.x[2 * (size_t)n];
    xa[1] = m.x[2 * (size_t)n + 1];
#pragma unroll
    for (int j = 0; j < 8; ++j) ga[j] = grad[8 * (size_t)n + j];
  }
  if (k < k1) load_face(k);
  for (; k < k1; k += 2) {
    const int ent = ent_n;
    const bool isR = ent < 0;
    const double nx = geo_n[0], ny = geo_n[1], len = geo_n[2];
    double qb[4], gb[8], xb[2], fc[2];
    const double phib = phib_n;
#pragma unroll
    for (int j = 0; j < 4; ++j) qb[j] = qb_n[j];
#pragma unroll
    for (int j = 0; j < 8; ++j) gb[j] = gb_n[j];
    xb[0] = xb_n[0]; xb[1] = xb_n[1];
    fc[0] = fc_n[0]; fc[1] = fc_n[1];
    if (k + 2 < k1) load_face(k + 2);
    PrimW Pb = prim_w(qb, g1);
    // this lane's contribution to block (n, other)
    float off[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) off[c] = 0.0f;
    if (len != 0.0) {
      double qa[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) qa[j] = nQ[4 * (size_t)n + j];
      PrimW Pa = Pn;
      if (recon == 1) {
        // reconstruct.C:181-205 at the subedge's face point, for both ends
        const double dxa0 = fc[0] - xa[0], dxa1 = fc[1] - xa[1];
        const double dxb0 = fc[0] - xb[0], dxb1 = fc[1] - xb[1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          qa[j] += phia * ga[2 * j] * dxa0 + phia * ga[2 * j + 1] * dxa1;
          qb[j] += phib * gb[2 * j] * dxb0 + phib * gb[2 * j + 1] * dxb1;
        }
        Pa = prim_w(qa, g1);
        Pb = prim_w(qb, g1);
      }
      // eigenvalue bound of the node's own state on this face
      ev += (fabs(Pn.u * nx + Pn.v * ny) + cn) * len;
      // the flux out of the node's control volume through this subedge:
      // Roe_flux_centered with the node's face state on the left and the
      // outward normal (the subedge normal, reversed where the node is the
      // subedge's R): R[L] += f and R[R] -= f of the reference's scatter
      const double ox = isR ? -nx : nx, oy = isR ? -ny : ny;
      double f[4];
      RoeWave w;
      roe_flux_wave_prim(ox, oy, len, g1, qa, qb, Pa, Pb, f, &w);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += f[j];
      if (jac) {
        // len/2 M = len/2 |A_roe| in FP32 from the face's Roe quantities
        const float hl = (float)(0.5 * len);
        const float fnx = (float)ox, fny = (float)oy;
        const float ub = (float)w.ub, vb = (float)w.vb, hb = (float)w.hb, th = (float)w.th, hV2 = (float)w.hV2;
        const float hat = hl * (float)w.at, k1 = hl * (float)w.k1, k2 = hl * (float)w.k2, k4 = hl * (float)w.k4;
        const float P[4] = {gm1 * hV2, -gm1 * ub, -gm1 * vb, gm1};
        const float U[4] = {-th, fnx, fny, 0.0f};
        const float ra[4] = {1.0f, ub, vb, hb}, rv[4] = {0.0f, fnx, fny, th};
        float w1[4], w2[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          w1[j] = k1 * P[j] + k2 * U[j];
          w2[j] = k2 * P[j] + k4 * U[j];
        }
        float hM[16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) hM[4 * i + j] = (i == j ? hat : 0.0f) + ra[i] * w1[j] + rv[i] * w2[j];
        float A[16];
        flux_jacobian_f32(fnx, fny, gf, (float)Pa.u, (float)Pa.v, (float)(qa[3] * Pa.ir), A);
#pragma unroll
        for (int c = 0; c < 16; ++c) dg[c] += hl * A[c] + hM[c];
        flux_jacobian_f32(fnx, fny, gf, (float)Pb.u, (float)Pb.v, (float)(qb[3] * Pb.ir), A);
#pragma unroll
        for (int c = 0; c < 16; ++c) off[c] = hl * A[c] - hM[c];
      }
    }
    // block (n, other) from the edge's two subedges in lanes v and v ^ 1;
    // lane v writes rows 2 (v & 1) and 2 (v & 1) + 1
#pragma unroll
    for (int c = 0; c < 16; ++c) off[c] += quad_xor<1>(off[c]);
    if (kOut == 0) {
      float row[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) row[j] = sub ? off[8 + j] : off[j];
      float4* blk = reinterpret_cast<float4*>(s.LHS32 + 16 * (size_t)m.einc_slot[k]) + 2 * sub;
      blk[0] = make_float4(row[0], row[1], row[2], row[3]);
      blk[1] = make_float4(row[4], row[5], row[6], row[7]);
    } else if (sub == 0) {
      // the whole block into the quad's LDS rows (both lanes hold it)
      float4* blk =
          reinterpret_cast<float4*>(s_rows + (threadIdx.x >> 2) * fused_row_stride(m.maxdeg) + 16 * (k - k0));
#pragma unroll
      for (int r = 0; r < 4; ++r) blk[r] = make_float4(off[4 * r], off[4 * r + 1], off[4 * r + 2], off[4 * r + 3]);
    }
  }
  // node sums: fixed butterfly over the quad
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] += quad_xor<1>(acc[j]);
  ev += quad_xor<1>(ev);
#pragma unroll
  for (int c = 0; c < 16; ++c) dg[c] += quad_xor<1>(dg[c]);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] += quad_xor<2>(acc[j]);
  ev += quad_xor<2>(ev);
#pragma unroll
  for (int c = 0; c < 16; ++c) dg[c] += quad_xor<2>(dg[c]);
  double Rv = v == 0 ? acc[0] : v == 1 ? acc[1] : v == 2 ? acc[2] : acc[3];
  for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
    const int b = m.nbe_list[k];
    const double bx = m.bedge_n[3 * (size_t)b], by = m.bedge_n[3 * (size_t)b + 1], bl = m.bedge_n[3 * (size_t)b + 2];
    ev += (fabs(Pn.u * bx + Pn.v * by) + cn) * bl;
    Rv += s.bflux[4 * (size_t)b + v];
  }
  const double V = m.cv_area[n];
  const double dtn = p.cfl * (V / ev);
  if (v == 0) {
    if (!(dtn > 0.0) || !isfinite(dtn) || !isfinite(ev)) atomicOr(s.err + inst, 2);
    s.ev_con[n] = ev;
    s.dt[n] = dtn;
  }
  const double vd = V / dtn;
  double D[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) D[c] = (c == v) ? vd : 0.0;
  if (jac) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      D[c] += (double)(v == 0 ? dg[c] : v == 1 ? dg[4 + c] : v == 2 ? dg[8 + c] : dg[12 + c]);
    for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
      const double* B = s.bjac + 16 * (size_t)m.nbe_list[k] + 4 * v;
#pragma unroll
      for (int c = 0; c < 4; ++c) D[c] += B[c];
    }
  }
  const float4 dv = make_float4((float)D[0], (float)D[1], (float)D[2], (float)D[3]);
  *reinterpret_cast<float4*>(s.D32 + 16 * (size_t)t + 4 * v) = dv;
  s.R[4 * (size_t)n + v] = Rv;
  s.RHS[4 * (size_t)n + v] = -Rv;
  const float bv = (float)(-Rv);
  if (kOut == 0) {
    s.RHS32[4 * (size_t)t + v] = bv;
    return;
  }
  // the stored diagonal block in every lane of the quad, and its inverse
  float A[16];
  {
    const float dr[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      A[c] = quad_bcast<0>(dr[c]);
      A[4 + c] = quad_bcast<1>(dr[c]);
      A[8 + c] = quad_bcast<2>(dr[c]);
      A[12 + c] = quad_bcast<3>(dr[c]);
    }
  }
  double Di[16], di[4];
  inv4_pivoted(A, Di);
#pragma unroll
  for (int i = 0; i < 4; ++i) di[i] = v == 0 ? Di[i] : v == 1 ? Di[4 + i] : v == 2 ? Di[8 + i] : Di[12 + i];
  const float4 rb = make_float4(quad_bcast<0>(bv), quad_bcast<1>(bv), quad_bcast<2>(bv), quad_bcast<3>(bv));
  s.RHS32[4 * (size_t)t + v] = scale_rhs_entry(di, rb);
  s.dQ32[4 * (size_t)t + v] = 0.0f;
  // the quad's staged blocks were written by its other lanes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float* rows = s_rows + (threadIdx.x >> 2) * fused_row_stride(m.maxdeg);
  bool ok = true;
  for (int e = 0; e < k1 - k0; ++e) {
    float b[16];
    const float4* B = reinterpret_cast<const float4*>(rows + 16 * e);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float4 q = B[r];
      b[4 * r] = q.x; b[4 * r + 1] = q.y; b[4 * r + 2] = q.z; b[4 * r + 3] = q.w;
    }
    const float4 o = scale_block_row(di, b);
    const size_t slot = (size_t)m.einc_slot[k0 + e];
    if (kOut == 1) {
      *reinterpret_cast<float4*>(s.LHS32 + 16 * slot + 4 * v) = o;
    } else {
      const unsigned int h0 = half_bits(o.x, ok), h1 = half_bits(o.y, ok), h2 = half_bits(o.z, ok),
                         h3 = half_bits(o.w, ok);
      *reinterpret_cast<uint2*>(s.LHS16 + 16 * slot + 4 * v) = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
    }
  }
  if (kOut == 2 && !ok) atomicOr(s.err + inst, 4);
}

// One thread per row (color position t): the diagonal block's inverse Di
// (pivoted, FP64).  Unscaled form (kScale false): Di is stored in FP32 for
// the sweep to apply.  Scaled form: the row's sweep operands are
// pre-multiplied by it in place instead, each rounded once to FP32 from
// FP64 products: every off-diagonal block A -> Di A and the right-hand side
// b -> Di b, so a sweep needs no diagonal: dQ_row = Di b - sum (Di A) dQ_nb.
// The node pass rewrites every block and right-hand side in every Newton
// iteration, so the scaling always starts from fresh operands; the diagonal
// block itself is left as it is.  Both forms zero the row's dQ, the sweeps'
// starting guess.
template <bool kScale>
__global__ void __launch_bounds__(kMaxBlock) k_diag_lu32(DevMesh m, const InstParams* __restrict__ ip,
                                                      float* __restrict__ LHS32, const float* __restrict__ D32,
                                                      float* __restrict__ LU32, float* __restrict__ RHS32,
                                                      float* __restrict__ dQ32) {
  int t = work_elem(m.pos_chunks, m.n_pos_chunks, xcd_block(m.xcd_run) * blockDim.x + threadIdx.x, m.nn);
  if (t >= m.nn) return;
  int n = m.color_rows[t];
  if (!ip[inst_of(m, m.pos_inst, t)].active) return;
  float A[16];
  double Di[16];
  const float4* D = reinterpret_cast<const float4*>(D32 + 16 * (size_t)t);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float4 r = D[k];
    A[4 * k] = r.x; A[4 * k + 1] = r.y; A[4 * k + 2] = r.z; A[4 * k + 3] = r.w;
  }
  inv4_pivoted(A, Di);
  *reinterpret_cast<float4*>(dQ32 + 4 * (size_t)t) = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!kScale) {
    float4* o = reinterpret_cast<float4*>(LU32 + 16 * (size_t)t);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = make_float4((float)Di[4 * k], (float)Di[4 * k + 1], (float)Di[4 * k + 2], (float)Di[4 * k + 3]);
    return;
  }
  const int deg = m.pdeg[t];
  const size_t base = (size_t)m.sbase[t >> 4] + (t & 15);
  for (int e = 0; e < deg; ++e) {
    const size_t k = base + 16 * (size_t)e;
    const float4* B = reinterpret_cast<const float4*>(LHS32 + 16 * k);
    float b[16];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float4 q = B[r];
      b[4 * r] = q.x; b[4 * r + 1] = q.y; b[4 * r + 2] = q.z; b[4 * r + 3] = q.w;
    }
    float4* O = reinterpret_cast<float4*>(LHS32 + 16 * k);
#pragma unroll
    for (int r = 0; r < 4; ++r) O[r] = scale_block_row(Di + 4 * r, b);
  }
  float4* R = reinterpret_cast<float4*>(RHS32 + 4 * (size_t)t);
  const float4 rb = *R;
  *R = make_float4(scale_rhs_entry(Di, rb), scale_rhs_entry(Di + 4, rb), scale_rhs_entry(Di + 8, rb),
                   scale_rhs_entry(Di + 12, rb));
}

// Multicolor sweeps of the mixed path over positions [row_begin,row_end) of
// the color order, 4 lanes per row (lane v: component v).  The sweep vectors
// (right-hand side, dQ) are indexed by color position and the column indices
// of the blocks are neighbour positions, so a wave's right-hand side reads
// and dQ writes are contiguous.  A row needs two dependent memory round
// trips: first its degree, activity byte, right-hand side and up to kPre
// neighbour positions (with a uniform slice width the slot base is
// arithmetic; the column array is padded so these loads need no bounds
// test), then its blocks and the neighbours' dQ.
__device__ __forceinline__ size_t gs_slot_base(const DevMesh& m, int t) {
  return m.slice_w ? (size_t)16 * m.slice_w * (size_t)(t >> 4) + (t & 15) : (size_t)m.sbase[t >> 4] + (t & 15);
}

// The row of a 4-lane sweep thread: the launch range from row_begin, or,
// with a work list of nchunks chunks, row (tid / 4) % 16 of the wave's chunk
// (one wave per chunk; its first position is wave-uniform, a scalar load);
// row_end (no row) past the last chunk.
__device__ __forceinline__ int gs_row(const int* __restrict__ chunks, int nchunks, int row_begin, int row_end,
                                      int tid) {
  if (!chunks) return row_begin + (tid >> 2);
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (w >= nchunks) return row_end;
  return chunks[w] + ((tid >> 2) & (kGsChunk - 1));
}

// FP32 blocks.  Unscaled form: r = b - sum A dQ_nb, then lane v applies row
// v of the stored diagonal inverse to the shuffled r.  Scaled form (kScale):
// the blocks and right-hand side already carry the inverse (k_diag_lu32<true>
// or the fused node pass), so dQ_row[v] = c[v] - sum (Di A)[v,:] dQ_nb.
// Neighbour dQ of a sweep row: kDpp false, every lane loads the neighbour's
// 4 components (16 bytes, the same address in all 4 lanes of the quad);
// kDpp true, lane v loads component v and the quad exchanges them with DPP
// broadcasts (a quarter of the load work; the same values, so the same
// bits).  Every lane of the quad must be active.
template <bool kDpp>
__device__ __forceinline__ float4 gs_neighbour(const float* __restrict__ dQ32, int nb, int v, float own) {
  if (!kDpp) return *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)nb);
  (void)v;
  return make_float4(quad_bcast<0>(own), quad_bcast<1>(own), quad_bcast<2>(own), quad_bcast<3>(own));
}

template <int kPre, bool kScale, bool kDpp>
__global__ void __launch_bounds__(kMaxBlock) k_gs_color32(DevMesh m, const unsigned char* __restrict__ pos_act,
                                                       const float* __restrict__ LHS32, const float* __restrict__ LU32,
                                                       const float* __restrict__ RHS32, float* __restrict__ dQ32,
                                                       const int* __restrict__ chunks, int nchunks, int row_begin,
                                                       int row_end) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  int t = gs_row(chunks, nchunks, row_begin, row_end, tid), v = tid & 3;
  if (t >= row_end) return;
  const size_t base = gs_slot_base(m, t);
  const int deg = m.pdeg[t];
  const int act = pos_act[t];
  float r = RHS32[4 * (size_t)t + v];
  int nb[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) nb[i] = m.ja[base + 16 * i];
  if (!act) return;
  float4 A[kPre], q[kPre];
  float qv[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    if (i < deg) {
      A[i] = *reinterpret_cast<const float4*>(LHS32 + 16 * (base + 16 * i) + 4 * v);
      if (kDpp)
        qv[i] = dQ32[4 * (size_t)nb[i] + v];
      else
        q[i] = *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)nb[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < kPre; ++i)
    if (i < deg) {
      const float4 b = kDpp ? gs_neighbour<true>(dQ32, 0, v, qv[i]) : q[i];
      r -= A[i].x * b.x + A[i].y * b.y + A[i].z * b.z + A[i].w * b.w;
    }
  for (int e = kPre; e < deg; ++e) {
    size_t k = base + 16 * (size_t)e;
    float4 a = *reinterpret_cast<const float4*>(LHS32 + 16 * k + 4 * v);
    const int nbk = m.ja[k];
    const float4 b = gs_neighbour<kDpp>(dQ32, nbk, v, kDpp ? dQ32[4 * (size_t)nbk + v] : 0.0f);
    r -= a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  }
  if (kScale) {
    dQ32[4 * (size_t)t + v] = r;
    return;
  }
  float b0 = __shfl(r, 0, 4), b1 = __shfl(r, 1, 4), b2 = __shfl(r, 2, 4), b3 = __shfl(r, 3, 4);
  // lane v: row v of the diagonal block's inverse
  const float4 Di = *reinterpret_cast<const float4*>(LU32 + 16 * (size_t)t + 4 * v);
  dQ32[4 * (size_t)t + v] = Di.x * b0 + Di.y * b1 + Di.z * b2 + Di.w * b3;
}

// The pre-multiplied sweep over FP16 operands: lane v reads row v of each
// off-diagonal block as 4 FP16 values (8 bytes; a wave reads block e of its
// 16 rows as one contiguous 512-byte run) and accumulates in FP32,
// dQ_row[v] = c[v] - sum (Di A)[v,:] dQ_nb.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
template <int kPre, bool kDpp>
__global__ void __launch_bounds__(kMaxBlock) k_gs_color16(DevMesh m, const unsigned char* __restrict__ pos_act,
                                                       const unsigned short* __restrict__ LHS16,
                                                       const float* __restrict__ RHS32, float* __restrict__ dQ32,
                                                       const int* __restrict__ chunks, int nchunks, int row_begin,
                                                       int row_end) {
  int tid = xcd_block(m.xcd_run) * blockDim.x + threadIdx.x;
  int t = gs_row(chunks, nchunks, row_begin, row_end, tid), v = tid & 3;
  if (t >= row_end) return;
  const size_t base = gs_slot_base(m, t);
  const int deg = m.pdeg[t];
  const int act = pos_act[t];
  float r = RHS32[4 * (size_t)t + v];
  int nb[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) nb[i] = m.ja[base + 16 * i];
  if (!act) return;
  half4_t A[kPre];
  float4 q[kPre];
  float qv[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    if (i < deg) {
      A[i] = *reinterpret_cast<const half4_t*>(LHS16 + 16 * (base + 16 * i) + 4 * v);
      if (kDpp)
        qv[i] = dQ32[4 * (size_t)nb[i] + v];
      else
        q[i] = *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)nb[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < kPre; ++i)
    if (i < deg) {
      const float4 b = kDpp ? gs_neighbour<true>(dQ32, 0, v, qv[i]) : q[i];
      r -= (float)A[i].x * b.x + (float)A[i].y * b.y + (float)A[i].z * b.z + (float)A[i].w * b.w;
    }
  for (int e = kPre; e < deg; ++e) {
    size_t k = base + 16 * (size_t)e;
    const half4_t a = *reinterpret_cast<const half4_t*>(LHS16 + 16 * k + 4 * v);
    const int nbk = m.ja[k];
    const float4 b = gs_neighbour<kDpp>(dQ32, nbk, v, kDpp ? dQ32[4 * (size_t)nbk + v] : 0.0f);
    r -= (float)a.x * b.x + (float)a.y * b.y + (float)a.z * b.z + (float)a.w * b.w;
  }
  dQ32[4 * (size_t)t + v] = r;
}

// The activity byte of every color position (its instance takes part in
// this Newton iteration), read by the sweeps with the row's other data
__global__ void __launch_bounds__(kBlock) k_pos_active(DevMesh m, const InstParams* __restrict__ ip,
                                                       unsigned char* __restrict__ pos_act) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.nn) return;
  pos_act[t] = ip[inst_of(m, m.pos_inst, t)].active ? 1 : 0;
}

// nQ += dQ (FP64 state, FP32 correction) and the FP64 norms' inputs; one
// thread per node (dQ32 is indexed by color position)
__global__ void __launch_bounds__(kBlock) k_update_mixed(DevMesh m, const InstParams* __restrict__ ip,
                                                         double* __restrict__ nQ, const float* __restrict__ dQ32,
                                                         double* __restrict__ dQ) {
  const int n = work_elem(m.node_chunks, m.n_node_chunks, blockIdx.x * blockDim.x + threadIdx.x, m.nn);
  if (n >= m.nn) return;
  if (!ip[inst_of(m, m.inst_node, n)].active) return;
  const float4 d = *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)(m.iau[n] - (m.nnz - m.nn)));
  double2* q = reinterpret_cast<double2*>(nQ + 4 * (size_t)n);
  double2* o = reinterpret_cast<double2*>(dQ + 4 * (size_t)n);
  const double2 d01 = make_double2((double)d.x, (double)d.y), d23 = make_double2((double)d.z, (double)d.w);
  double2 q01 = q[0], q23 = q[1];
  q01.x += d01.x; q01.y += d01.y; q23.x += d23.x; q23.y += d23.y;
  o[0] = d01;
  o[1] = d23;
  q[0] = q01;
  q[1] = q23;
}

namespace {
// LDS of one workgroup of the fused node pass
constexpr size_t kFusedLds = 65536;
size_t fused_lds_bytes(const DevMesh& m, int block) {
  return (size_t)(block / 4) * fused_row_stride(m.maxdeg) * sizeof(float);
}
// workgroup size of the fused node pass: the tuned one, halved until a
// workgroup's staged rows fit; 0 if even 16 rows do not
int fused_block(const DevMesh& m) {
  int b = m.tune.node_block;
  while (b > 64 && fused_lds_bytes(m, b) > kFusedLds) b /= 2;
  return fused_lds_bytes(m, b) <= kFusedLds ? b : 0;
}
}  // namespace

bool resid_jac32_fused(const DevMesh& m, const DevState& s) {
  if (s.LHS16) return true;
  return m.tune.gs32_scaled && m.tune.node_fused && fused_block(m) > 0;
}
bool resid_jac32_fused_fits(const DevMesh& m) { return fused_block(m) > 0; }

void launch_resid_jac32(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  if (!resid_jac32_fused(m, s)) {
    hipLaunchKernelGGL(k_resid_jac32<0>, dim3(nblocks(4 * nn, m.tune.node_block)), dim3(m.tune.node_block), 0, st,
                       m, s);
    return;
  }
  const int b = fused_block(m);
  const dim3 grid(nblocks(4 * nn, b)), block(b);
  if (s.LHS16)
    hipLaunchKernelGGL(k_resid_jac32<2>, grid, block, fused_lds_bytes(m, b), st, m, s);
  else
    hipLaunchKernelGGL(k_resid_jac32<1>, grid, block, fused_lds_bytes(m, b), st, m, s);
}
void launch_diag_lu32(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0 || resid_jac32_fused(m, s)) return;
  const int np = work_count(m.pos_chunks, m.n_pos_chunks, m.nn);
  if (np == 0) return;
  const dim3 grid(nblocks(np, m.tune.cell_block)), block(m.tune.cell_block);
  if (m.tune.gs32_scaled)
    hipLaunchKernelGGL(k_diag_lu32<true>, grid, block, 0, st, m, s.ip, s.LHS32, s.D32, s.LU32, s.RHS32, s.dQ32);
  else
    hipLaunchKernelGGL(k_diag_lu32<false>, grid, block, 0, st, m, s.ip, s.LHS32, s.D32, s.LU32, s.RHS32, s.dQ32);
}

namespace {
template <int kPre, bool kDpp>
void gs32_launch_form(const DevMesh& m, const DevState& s, dim3 grid, dim3 block, const int* ch, int nch, int b,
                      int e, hipStream_t st) {
  if (s.LHS16)
    hipLaunchKernelGGL((k_gs_color16<kPre, kDpp>), grid, block, 0, st, m, s.pos_act, s.LHS16, s.RHS32, s.dQ32, ch,
                       nch, b, e);
  else if (m.tune.gs32_scaled)
    hipLaunchKernelGGL((k_gs_color32<kPre, true, kDpp>), grid, block, 0, st, m, s.pos_act, s.LHS32, s.LU32, s.RHS32,
                       s.dQ32, ch, nch, b, e);
  else
    hipLaunchKernelGGL((k_gs_color32<kPre, false, kDpp>), grid, block, 0, st, m, s.pos_act, s.LHS32, s.LU32, s.RHS32,
                       s.dQ32, ch, nch, b, e);
}
template <int kPre>
void gs32_launch(const DevMesh& m, const DevState& s, dim3 grid, dim3 block, const int* ch, int nch, int b, int e,
                 hipStream_t st) {
  if (m.tune.gs32_dpp)
    gs32_launch_form<kPre, true>(m, s, grid, block, ch, nch, b, e, st);
  else
    gs32_launch_form<kPre, false>(m, s, grid, block, ch, nch, b, e, st);
}
}  // namespace

void launch_gs32_sweeps(const DevMesh& m, const DevState& s, const ColorRanges& cr, int subits, hipStream_t st) {
  if (m.nn > 0) hipLaunchKernelGGL(k_pos_active, dim3(nblocks(m.nn)), dim3(kBlock), 0, st, m, s.ip, s.pos_act);
  std::vector<GsLaunch> sched((size_t)subits * m.ncolor + 1);
  int ns = gs_schedule(m.ncolor, subits, sched.data());
  for (int i = 0; i < ns; ++i) {
    int c = sched[i].color;
    int b = cr.off[c], e = sched[i].backward ? b + cr.nback[c] : cr.off[c + 1];
    if (e <= b) continue;
    const int bs = m.tune.gs32_block;
    const int* ch = nullptr;
    int nch = 0, nthreads = 4 * (e - b);
    if (cr.chunks) {
      const int r = 2 * c + sched[i].backward;
      ch = cr.chunks + cr.chunk_off[r];
      nch = cr.chunk_off[r + 1] - cr.chunk_off[r];
      nthreads = 4 * kGsChunk * nch;
      if (nthreads == 0) continue;
    }
    const dim3 grid(nblocks(nthreads, bs));
    if (m.tune.gs32_pre <= 4)
      gs32_launch<4>(m, s, grid, dim3(bs), ch, nch, b, e, st);
    else if (m.tune.gs32_pre <= 6)
      gs32_launch<6>(m, s, grid, dim3(bs), ch, nch, b, e, st);
    else
      gs32_launch<8>(m, s, grid, dim3(bs), ch, nch, b, e, st);
  }
}
void launch_update_mixed(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  const int nn = work_count(m.node_chunks, m.n_node_chunks, m.nn);
  if (nn == 0) return;
  hipLaunchKernelGGL(k_update_mixed, dim3(nblocks(nn)), dim3(kBlock), 0, st, m, s.ip, s.nQ, s.dQ32, s.dQ);
}

}  // namespace fsk
