// Device-side math for the median-dual Euler path on gfx950.
//
// FP64 functions keep the reference's operation order term by term (the
// library is built with -ffp-contract=off) so that a face flux evaluated here
// equals the reference CPU value bit for bit.  That matters beyond the
// residual: the reference's Jacobian is a forward difference with eps=1e-8
// (jacobian.C:54-106), which amplifies any last-bit difference of the flux by
// 1e8.  FP32 functions (mixed precision) use a division-light form instead.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_face.hpp"  // per-face FP64 math shared with host code

namespace fsd {

// ---------------------------------------------------------------------------
// FP64, reference operation order
// ---------------------------------------------------------------------------

// 1/x for the mixed path's FP64 residual: the hardware reciprocal estimate
// refined by two Newton steps (about one ulp; the FP64 path keeps IEEE
// division for the reference's bits).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Primitive quantities of one state for the mixed path's FP64 residual:
// 1/rho, u, v, p and the total enthalpy per volume H = E + p.
struct PrimW {
  double ir, u, v, p, H;
};
__device__ __forceinline__ PrimW prim_w(const double* q, double g1) {
  PrimW s;
  s.ir = rcp_nr(q[0]);
  s.u = q[1] * s.ir;
  s.v = q[2] * s.ir;
  s.p = g1 * (q[3] - 0.5 * q[0] * (s.u * s.u + s.v * s.v));
  s.H = q[3] + s.p;
  return s;
}

// The Roe-averaged quantities of one face from which its dissipation matrix
// is rebuilt: |A_roe| = at I + ra (x) w1 + rb (x) w2 with ra = (1, ub, vb, hb),
// rb = (0, nx, ny, th), w1 = k1 P + k2 U, w2 = k2 P + k4 U, where
// P = g1 (hV2, -ub, -vb, 1) (dp = P.dQ, the pressure jump) and
// U = (-th, nx, ny, 0) (dun = U.dQ, rho times the normal-velocity jump).
struct RoeWave {
  double ub, vb, hb, th, hV2, at, k1, k2, k4;
};

// Roe_flux_centered (flux.C:299-541) for the mixed path's FP64 residual:
// f = 0.5 len (F(qL).n + F(qR).n - |A_roe| (qR - qL)) with the dissipation in
// closed form,
//   |A_roe| dQ = at dQ + ra alpha + rb beta,
//   alpha = k1 dp + k2 dun,  beta = k2 dp + k4 dun,
// where at = |th|, S = (|th+c| + |th-c|)/2, D = (|th+c| - |th-c|)/2,
// k4 = S - at, k1 = k4 / c^2, k2 = D / c.  This is T |Lambda| Tinv dQ of the
// reference (the entropy and shear waves move with |th|, the two acoustic
// waves add the rank-2 correction), with 2 reciprocals (rcp_nr) and 2 square
// roots per face besides the states' own 1/rho.  Same mathematics as roe_flux,
// different rounding: agrees with it to a few ulps of the flux magnitude, not
// bit for bit.  The Roe quantities go to *w for the FP32 Jacobian.
__device__ __forceinline__ void roe_flux_wave_prim(double nx, double ny, double len, double g1, const double* qL,
                                                   const double* qR, const PrimW& L, const PrimW& R, double* flux,
                                                   RoeWave* w) {
  const double thL = L.u * nx + L.v * ny, thR = R.u * nx + R.v * ny;
  const double rb = sqrt(qL[0] * qR[0]);
  const double bg = rb * rcp_nr(qL[0] + rb);
  const double hL = L.H * L.ir, hR = R.H * R.ir;
  const double ub = L.u + bg * (R.u - L.u), vb = L.v + bg * (R.v - L.v), hb = hL + bg * (hR - hL);
  const double th = ub * nx + vb * ny;
  const double hV2 = 0.5 * (ub * ub + vb * vb);
  const double c = sqrt(g1 * (hb - hV2));
  const double ic = rcp_nr(c);
  const double d0 = qR[0] - qL[0], d1 = qR[1] - qL[1], d2 = qR[2] - qL[2], d3 = qR[3] - qL[3];
  const double dp = g1 * (hV2 * d0 - ub * d1 - vb * d2 + d3);
  const double dun = nx * d1 + ny * d2 - th * d0;
  const double at = fabs(th), lp = fabs(th + c), lm = fabs(th - c);
  const double k4 = 0.5 * (lp + lm) - at, k2 = 0.5 * (lp - lm) * ic, k1 = k4 * (ic * ic);
  const double al = k1 * dp + k2 * dun, be = k2 * dp + k4 * dun;
  const double D0 = at * d0 + al;
  const double D1 = at * d1 + ub * al + nx * be;
  const double D2 = at * d2 + vb * al + ny * be;
  const double D3 = at * d3 + hb * al + th * be;
  const double hl = 0.5 * len;
  flux[0] = hl * (qL[0] * thL + qR[0] * thR - D0);
  flux[1] = hl * (qL[1] * thL + L.p * nx + qR[1] * thR + R.p * nx - D1);
  flux[2] = hl * (qL[2] * thL + L.p * ny + qR[2] * thR + R.p * ny - D2);
  flux[3] = hl * (L.H * thL + R.H * thR - D3);
  if (w) {
    w->ub = ub; w->vb = vb; w->hb = hb; w->th = th; w->hV2 = hV2;
    w->at = at; w->k1 = k1; w->k2 = k2; w->k4 = k4;
  }
}
__device__ __forceinline__ void roe_flux_wave(double nx, double ny, double len, double gamma, const double* qL,
                                              const double* qR, double* flux) {
  const double g1 = gamma - 1.0;
  roe_flux_wave_prim(nx, ny, len, g1, qL, qR, prim_w(qL, g1), prim_w(qR, g1), flux, nullptr);
}

// unpivoted 4x4 LU and substitution (linear_algebra.C:571-638)
__device__ __forceinline__ void lu4(double* A) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int i = k + 1; i < 4; i++) A[i * 4 + k] = A[i * 4 + k] / A[k * 4 + k];
#pragma unroll
    for (int j = k + 1; j < 4; j++)
#pragma unroll
      for (int i = k + 1; i < 4; i++) A[i * 4 + j] = A[i * 4 + j] - A[i * 4 + k] * A[k * 4 + j];
  }
}
template <typename T>
__device__ __forceinline__ void lu_solve4(const T* A, const T* b, T* x) {
  T y0 = b[0];
  T y1 = b[1] - A[4] * y0;
  T y2 = b[2] - A[8] * y0 - A[9] * y1;
  T y3 = b[3] - A[12] * y0 - A[13] * y1 - A[14] * y2;
  x[3] = y3 / A[15];
  x[2] = (y2 - A[11] * x[3]) / A[10];
  x[1] = (y1 - A[6] * x[2] - A[7] * x[3]) / A[5];
  x[0] = (y0 - A[1] * x[1] - A[2] * x[2] - A[3] * x[3]) / A[0];
}

}  // namespace fsd
