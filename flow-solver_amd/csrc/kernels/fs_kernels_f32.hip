// Kernels of the mixed-precision path (gfx950, wave64).
//
// The outer loop stays FP64: the node kernel k_node_mixed evaluates the FP64
// residual (characteristic-wave Roe flux at the FP64 face states), ev_con and
// dt; the limiter, gradient, derivatives, norms, convergence test and state
// update are the FP64 kernels.  Inside it the Jacobian blocks and the
// Gauss-Seidel sweeps run in FP32: the analytic Roe Jacobian of
// compute_approximate_jacobian_roe (jacobian.C:301-658) at the same
// reconstructed face states the reference differentiates around, FP32 block
// storage (64 B per 4x4 block), diagonal blocks inverted in FP64 with
// pivoting and stored FP32, and FP32 sweeps.  The fixed point is the FP64
// residual's, so converged states agree with the FP64 path to the FP64
// stopping tolerance.
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

// Roe matrix T|Lambda|Tinv in FP32 with shared reciprocals
__device__ __forceinline__ void roe_matrix_f32(float nx, float ny, float g, const float* qL, const float* qR,
                                               float* Rm) {
  float irL = 1.0f / qL[0], irR = 1.0f / qR[0];
  float uL = qL[1] * irL, vL = qL[2] * irL, uR = qR[1] * irR, vR = qR[2] * irR;
  float gm1 = g - 1.0f;
  float pL = gm1 * (qL[3] - 0.5f * qL[0] * (uL * uL + vL * vL));
  float pR = gm1 * (qR[3] - 0.5f * qR[0] * (uR * uR + vR * vR));
  float hL = (qL[3] + pL) * irL, hR = (qR[3] + pR) * irR;
  float rb = sqrtf(qL[0] * qR[0]);
  float bg = rb / (qL[0] + rb);
  float ub = uL + bg * (uR - uL), vb = vL + bg * (vR - vL), hb = hL + bg * (hR - hL);
  float th = ub * nx + vb * ny;
  float V2 = ub * ub + vb * vb;
  float c = sqrtf(gm1 * (hb - 0.5f * V2));
  float ic = 1.0f / c, ic2 = ic * ic;
  float lam[4] = {fabsf(th), fabsf(th), fabsf(th + c), fabsf(th - c)};
  float T[16], Ti[16];
  T[0] = 1.0f; T[1] = 0.0f; T[2] = ic2; T[3] = ic2;
  T[4] = ub; T[5] = -ny; T[6] = ub * ic2 + nx * ic; T[7] = ub * ic2 - nx * ic;
  T[8] = vb; T[9] = nx; T[10] = vb * ic2 + ny * ic; T[11] = vb * ic2 - ny * ic;
  T[12] = 0.5f * V2; T[13] = -ub * ny + vb * nx;
  float igm1 = 1.0f / gm1;
  T[14] = igm1 + th * ic + 0.5f * V2 * ic2;
  T[15] = igm1 - th * ic + 0.5f * V2 * ic2;
  Ti[0] = 1.0f - gm1 * 0.5f * V2 * ic2; Ti[1] = gm1 * ub * ic2; Ti[2] = gm1 * vb * ic2; Ti[3] = -gm1 * ic2;
  Ti[4] = ub * ny - vb * nx; Ti[5] = -ny; Ti[6] = nx; Ti[7] = 0.0f;
  Ti[8] = 0.5f * (-c * th + gm1 * 0.5f * V2); Ti[9] = 0.5f * (c * nx - gm1 * ub);
  Ti[10] = 0.5f * (c * ny - gm1 * vb); Ti[11] = 0.5f * gm1;
  Ti[12] = 0.5f * (c * th + gm1 * 0.5f * V2); Ti[13] = -0.5f * (c * nx + gm1 * ub);
  Ti[14] = -0.5f * (c * ny + gm1 * vb); Ti[15] = 0.5f * gm1;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float a = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) a += T[i * 4 + k] * (lam[k] * Ti[k * 4 + j]);
      Rm[i * 4 + j] = a;
    }
}

// Euler flux Jacobian dF.n/dq (jacobian.C:561-582)
__device__ __forceinline__ void flux_jacobian_f32(float nx, float ny, float g, const float* q, float* d) {
  float ir = 1.0f / q[0];
  float u = q[1] * ir, v = q[2] * ir, E = q[3];
  float th = u * nx + v * ny;
  float ph = 0.5f * (g - 1.0f) * (u * u + v * v);
  float gE = g * E * ir;
  d[0] = 0.0f; d[1] = nx; d[2] = ny; d[3] = 0.0f;
  d[4] = ph * nx - u * th; d[5] = (2.0f - g) * u * nx + th; d[6] = u * ny - (g - 1.0f) * v * nx; d[7] = (g - 1.0f) * nx;
  d[8] = ph * ny - v * th; d[9] = v * nx - (g - 1.0f) * u * ny; d[10] = th + (2.0f - g) * v * ny; d[11] = (g - 1.0f) * ny;
  d[12] = (2.0f * ph - gE) * th; d[13] = (1.0f - g) * u * th + nx * (gE - ph);
  d[14] = (1.0f - g) * v * th + ny * (gE - ph); d[15] = g * th;
}

// Edge-centric FP32 analytic Jacobian: for each of the edge's subedges,
// dF/dqL = 0.5(A(qL) + |A_roe|) len and dF/dqR = 0.5(A(qR) - |A_roe|) len at
// the (limited) reconstructed face states; block (L,R) += dF/dqR and block
// (R,L) -= dF/dqL (jacobian.C:612-658).  Zeros when jupdate is off.
__global__ void __launch_bounds__(kMaxBlock) k_edge_jac32(DevMesh m, DevState st) {
  int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const InstParams& ip = st.ip[inst_of(m, m.inst_edge, e)];
  if (!ip.active) return;
  const int recon = ip.recon_flux, jupdate = ip.jac;
  const float gamma = (float)ip.gamma;
  const double* __restrict__ nQ = st.nQ;
  const double* __restrict__ grad = st.grad;
  const double* __restrict__ phi = st.phi;
  float* __restrict__ LHS32 = st.LHS32;
  float offLR[16], offRL[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) { offLR[k] = 0.0f; offRL[k] = 0.0f; }
  if (jupdate) {
    int L = m.edges[2 * e], R = m.edges[2 * e + 1];
    float qL0[4], qR0[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      qL0[j] = (float)nQ[4 * (size_t)L + j];
      qR0[j] = (float)nQ[4 * (size_t)R + j];
    }
    float gL[8], gR[8], phL = 0.f, phR = 0.f, xL0 = 0.f, xL1 = 0.f, xR0 = 0.f, xR1 = 0.f;
    if (recon == 1) {
#pragma unroll
      for (int k = 0; k < 8; ++k) { gL[k] = (float)grad[8 * (size_t)L + k]; gR[k] = (float)grad[8 * (size_t)R + k]; }
      phL = (float)phi[L];
      phR = (float)phi[R];
      xL0 = (float)m.x[2 * (size_t)L]; xL1 = (float)m.x[2 * (size_t)L + 1];
      xR0 = (float)m.x[2 * (size_t)R]; xR1 = (float)m.x[2 * (size_t)R + 1];
    }
    for (int t = 0; t < 2; ++t) {
      int s = m.edge_sub[2 * e + t];
      if (s < 0) break;
      float nx = (float)m.sub_n[3 * (size_t)s], ny = (float)m.sub_n[3 * (size_t)s + 1];
      float len = (float)m.sub_n[3 * (size_t)s + 2];
      float ql[4], qr[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) { ql[j] = qL0[j]; qr[j] = qR0[j]; }
      if (recon == 1) {
        size_t ki = (size_t)m.sub_inc[2 * (size_t)s];
        float fc0 = (float)m.inc_fc[2 * ki], fc1 = (float)m.inc_fc[2 * ki + 1];
        float dl0 = fc0 - xL0, dl1 = fc1 - xL1, dr0 = fc0 - xR0, dr1 = fc1 - xR1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ql[j] += phL * (gL[2 * j] * dl0 + gL[2 * j + 1] * dl1);
          qr[j] += phR * (gR[2 * j] * dr0 + gR[2 * j + 1] * dr1);
        }
      }
      float Rm[16], A[16];
      roe_matrix_f32(nx, ny, gamma, ql, qr, Rm);
      float hl = 0.5f * len;
      flux_jacobian_f32(nx, ny, gamma, ql, A);
#pragma unroll
      for (int k = 0; k < 16; ++k) offRL[k] -= hl * (A[k] + Rm[k]);
      flux_jacobian_f32(nx, ny, gamma, qr, A);
#pragma unroll
      for (int k = 0; k < 16; ++k) offLR[k] += hl * (A[k] - Rm[k]);
    }
  }
  float4* blr = reinterpret_cast<float4*>(LHS32 + 16 * (size_t)m.edge_slot[2 * e]);
  float4* brl = reinterpret_cast<float4*>(LHS32 + 16 * (size_t)m.edge_slot[2 * e + 1]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    blr[k] = make_float4(offLR[4 * k], offLR[4 * k + 1], offLR[4 * k + 2], offLR[4 * k + 3]);
    brl[k] = make_float4(offRL[4 * k], offRL[4 * k + 1], offRL[4 * k + 2], offRL[4 * k + 3]);
  }
}

// Node pass of the mixed step: FP64 ev_con/dt (same order as the FP64 path),
// the FP64 interior residual, FP32 diagonal block V/dt I - sum of transposed
// off-diagonals + boundary blocks, FP32 right-hand side -R.
//
// The residual is accumulated here rather than through a per-subedge flux
// buffer: lane v of the node's 4 evaluates the characteristic-wave Roe flux
// of the node's subedge incidences v, v+4, ... (the flux of the subedge with
// the reference's L/R orientation and reconstruction, +f where the node is
// L and -f where it is R), then the 4 partial sums are combined by a fixed
// butterfly.  Each subedge is evaluated at both of its nodes (twice the
// flux arithmetic, none of the buffer's 64 bytes of writes and reads per
// subedge); the sum order depends only on the node's own incidence list, so
// an instance's residual is the same alone and in a batch.
__global__ void __launch_bounds__(kMaxBlock) k_node_mixed(DevMesh m, DevState s) {
  int tid = xcd_block() * blockDim.x + threadIdx.x;
  int n = tid >> 2, v = tid & 3;
  if (n >= m.nn) return;
  const int inst = inst_of(m, m.inst_node, n);
  const InstParams& p = s.ip[inst];
  if (!p.active) return;
  const int jac = p.jac;
  double q[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) q[j] = s.nQ[4 * (size_t)n + j];
  double Rv = 0.0, ev = 0.0;
  const int k0 = m.nsub_off[n], k1 = m.nsub_off[n + 1];
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    const double* nr = m.inc_n + 4 * (size_t)k;
    ev += (max_ev(q, nr[0], nr[1], p.gamma) * nr[2]);
  }
  {
    const int recon = p.recon_flux;
    const double* __restrict__ grad = s.grad;
    double gs[8], phis = 0.0, xs0 = 0.0, xs1 = 0.0;
    if (recon == 1) {
#pragma unroll
      for (int c = 0; c < 8; ++c) gs[c] = grad[8 * (size_t)n + c];
      phis = s.phi[n];
      xs0 = m.x[2 * (size_t)n];
      xs1 = m.x[2 * (size_t)n + 1];
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = k0 + v; k < k1; k += 4) {
      const int ent = m.inc_sother[k];
      const int o = ent & 0x7fffffff;
      const bool isR = ent < 0;
      const double* nr = m.inc_n + 4 * (size_t)k;
      double qs[4], qo[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        qs[j] = q[j];
        qo[j] = s.nQ[4 * (size_t)o + j];
      }
      if (recon == 1) {
        // reconstruct.C:181-205 at the subedge's face point, for both ends
        const double fc0 = m.inc_fc[2 * (size_t)k], fc1 = m.inc_fc[2 * (size_t)k + 1];
        const double dxs0 = fc0 - xs0, dxs1 = fc1 - xs1;
        const double dxo0 = fc0 - m.x[2 * (size_t)o], dxo1 = fc1 - m.x[2 * (size_t)o + 1];
        const double phio = s.phi[o];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          qs[j] += (phis * gs[2 * j] * dxs0 + phis * gs[2 * j + 1] * dxs1);
          qo[j] += (phio * grad[8 * (size_t)o + 2 * j] * dxo0 + phio * grad[8 * (size_t)o + 2 * j + 1] * dxo1);
        }
      }
      double ql[4], qr[4], f[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ql[j] = isR ? qo[j] : qs[j];
        qr[j] = isR ? qs[j] : qo[j];
      }
      roe_flux_wave(nr[0], nr[1], nr[2], p.gamma, ql, qr, f);
      // R[L] += f, R[R] -= f; adding -f is the same operation as subtracting f
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += isR ? -f[j] : f[j];
    }
#pragma unroll
    for (int off = 1; off <= 2; off <<= 1)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += __shfl_xor(acc[j], off, 4);
    Rv = v == 0 ? acc[0] : v == 1 ? acc[1] : v == 2 ? acc[2] : acc[3];
  }
  for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
    int b = m.nbe_list[k];
    double nx = m.bedge_n[3 * (size_t)b], ny = m.bedge_n[3 * (size_t)b + 1], len = m.bedge_n[3 * (size_t)b + 2];
    ev += (max_ev(q, nx, ny, p.gamma) * len);
    Rv += s.bflux[4 * (size_t)b + v];
  }
  double V = m.cv_area[n];
  double dtn = p.cfl * (V / ev);
  if (v == 0) {
    if (!(dtn > 0.0) || !isfinite(dtn) || !isfinite(ev)) atomicOr(s.err + inst, 2);
    s.ev_con[n] = ev;
    s.dt[n] = dtn;
  }
  double vd = V / dtn;
  double D[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) D[c] = (c == v) ? vd : 0.0;
  if (jac) {
#pragma unroll 4
    for (int q = m.roff[n]; q < m.roff[n + 1]; ++q) {
      float4 B = *reinterpret_cast<const float4*>(s.LHS32 + 16 * (size_t)m.tslot[q] + 4 * v);
      D[0] -= B.x; D[1] -= B.y; D[2] -= B.z; D[3] -= B.w;
    }
    for (int k = m.nbe_off[n]; k < m.nbe_off[n + 1]; ++k) {
      const double* B = s.bjac + 16 * (size_t)m.nbe_list[k] + 4 * v;
#pragma unroll
      for (int c = 0; c < 4; ++c) D[c] += B[c];
    }
  }
  *reinterpret_cast<float4*>(s.LHS32 + 16 * (size_t)m.iau[n] + 4 * v) =
      make_float4((float)D[0], (float)D[1], (float)D[2], (float)D[3]);
  s.R[4 * (size_t)n + v] = Rv;
  s.RHS[4 * (size_t)n + v] = -Rv;
  s.RHS32[4 * (size_t)n + v] = (float)(-Rv);
}

// Inverse of a 4x4 block in FP64 by Gauss-Jordan elimination with partial
// pivoting, stored in FP32.  A factorization without pivoting rounded to FP32
// loses the block where its leading pivot is small against the other entries
// (the continuity row's diagonal is V/dt plus dissipation only, since the
// flux Jacobian's (0,0) entry is zero); the pivoted FP64 inverse is accurate
// to FP32 rounding of its entries.  Row selection uses compile-time indices
// only, so the augmented matrix stays in registers.
__device__ __forceinline__ void inv4_pivoted(const float* A, float* out) {
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
    for (int j = 0; j < 4; ++j) out[4 * i + j] = (float)a[i][4 + j];
}

__global__ void __launch_bounds__(kMaxBlock) k_diag_lu32(DevMesh m, const InstParams* __restrict__ ip,
                                                      const float* __restrict__ LHS32, float* __restrict__ LU32,
                                                      float* __restrict__ dQ32) {
  int t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= m.nn) return;
  int n = m.color_rows[t];
  if (!ip[inst_of(m, m.pos_inst, t)].active) return;
  float A[16], Ai[16];
  const float4* D = reinterpret_cast<const float4*>(LHS32 + 16 * (size_t)m.iau[n]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float4 r = D[k];
    A[4 * k] = r.x; A[4 * k + 1] = r.y; A[4 * k + 2] = r.z; A[4 * k + 3] = r.w;
  }
  inv4_pivoted(A, Ai);
  float4* o = reinterpret_cast<float4*>(LU32 + 16 * (size_t)t);
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = make_float4(Ai[4 * k], Ai[4 * k + 1], Ai[4 * k + 2], Ai[4 * k + 3]);
  *reinterpret_cast<float4*>(dQ32 + 4 * (size_t)n) = make_float4(0.f, 0.f, 0.f, 0.f);
}

// FP32 multicolor sweep over rows [row_begin,row_end) of the color order;
// 4 lanes per row (lane v: component v), quad-wide FP32 block rows of the
// sliced layout; the row's index, block and neighbour loads are issued
// before the arithmetic.
template <int kPre>
__global__ void __launch_bounds__(kMaxBlock) k_gs_color32(DevMesh m, const InstParams* __restrict__ ip,
                                                       const float* __restrict__ LHS32, const float* __restrict__ LU32,
                                                       const float* __restrict__ RHS32, float* __restrict__ dQ32,
                                                       int row_begin, int row_end) {
  int tid = xcd_block() * blockDim.x + threadIdx.x;
  int t = row_begin + (tid >> 2), v = tid & 3;
  if (t >= row_end) return;
  int j = m.color_rows[t];
  int inst = inst_of(m, m.pos_inst, t);
  int deg = m.pdeg[t];
  size_t base = (size_t)m.sbase[t >> 4] + (t & 15);
  float r = RHS32[4 * (size_t)j + v];
  if (!ip[inst].active) return;
  int nb[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) nb[i] = (i < deg) ? m.ja[base + 16 * i] : j;
  float4 A[kPre], q[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    if (i < deg) {
      A[i] = *reinterpret_cast<const float4*>(LHS32 + 16 * (base + 16 * i) + 4 * v);
      q[i] = *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)nb[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < kPre; ++i)
    if (i < deg) r -= A[i].x * q[i].x + A[i].y * q[i].y + A[i].z * q[i].z + A[i].w * q[i].w;
  for (int e = kPre; e < deg; ++e) {
    size_t k = base + 16 * (size_t)e;
    float4 a = *reinterpret_cast<const float4*>(LHS32 + 16 * k + 4 * v);
    float4 b = *reinterpret_cast<const float4*>(dQ32 + 4 * (size_t)m.ja[k]);
    r -= a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  }
  float b0 = __shfl(r, 0, 4), b1 = __shfl(r, 1, 4), b2 = __shfl(r, 2, 4), b3 = __shfl(r, 3, 4);
  // lane v: row v of the diagonal block's inverse
  const float4 Di = *reinterpret_cast<const float4*>(LU32 + 16 * (size_t)t + 4 * v);
  dQ32[4 * (size_t)j + v] = Di.x * b0 + Di.y * b1 + Di.z * b2 + Di.w * b3;
}

// nQ += dQ (FP64 state, FP32 correction) and the FP64 norms' inputs
__global__ void __launch_bounds__(kBlock) k_update_mixed(DevMesh m, const InstParams* __restrict__ ip,
                                                         double* __restrict__ nQ, const float* __restrict__ dQ32,
                                                         double* __restrict__ dQ) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * m.nn) return;
  if (!ip[inst_of(m, m.inst_node, i >> 2)].active) return;
  double d = (double)dQ32[i];
  dQ[i] = d;
  nQ[i] += d;
}

void launch_edge_jac32(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.ne == 0) return;
  hipLaunchKernelGGL(k_edge_jac32, dim3(nblocks(m.ne, m.tune.face_block)), dim3(m.tune.face_block), 0, st, m, s);
}
void launch_node_mixed(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  hipLaunchKernelGGL(k_node_mixed, dim3(nblocks(4 * m.nn, m.tune.node_block)), dim3(m.tune.node_block), 0, st, m, s);
}
void launch_diag_lu32(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  hipLaunchKernelGGL(k_diag_lu32, dim3(nblocks(m.nn, m.tune.cell_block)), dim3(m.tune.cell_block), 0, st, m, s.ip,
                     s.LHS32, s.LU32, s.dQ32);
}

void launch_gs32_sweeps(const DevMesh& m, const DevState& s, const ColorRanges& cr, int subits, hipStream_t st) {
  std::vector<GsLaunch> sched((size_t)subits * m.ncolor + 1);
  int ns = gs_schedule(m.ncolor, subits, sched.data());
  for (int i = 0; i < ns; ++i) {
    int c = sched[i].color;
    int b = cr.off[c], e = sched[i].backward ? b + cr.nback[c] : cr.off[c + 1];
    if (e <= b) continue;
    const int bs = m.tune.gs32_block;
    const dim3 grid(nblocks(4 * (e - b), bs));
    if (m.tune.gs32_pre <= 4)
      hipLaunchKernelGGL(k_gs_color32<4>, grid, dim3(bs), 0, st, m, s.ip, s.LHS32, s.LU32, s.RHS32, s.dQ32, b, e);
    else if (m.tune.gs32_pre <= 6)
      hipLaunchKernelGGL(k_gs_color32<6>, grid, dim3(bs), 0, st, m, s.ip, s.LHS32, s.LU32, s.RHS32, s.dQ32, b, e);
    else
      hipLaunchKernelGGL(k_gs_color32<8>, grid, dim3(bs), 0, st, m, s.ip, s.LHS32, s.LU32, s.RHS32, s.dQ32, b, e);
  }
}
void launch_update_mixed(const DevMesh& m, const DevState& s, hipStream_t st) {
  if (m.nn == 0) return;
  hipLaunchKernelGGL(k_update_mixed, dim3(nblocks(4 * m.nn)), dim3(kBlock), 0, st, m, s.ip, s.nQ, s.dQ32, s.dQ);
}

}  // namespace fsk
