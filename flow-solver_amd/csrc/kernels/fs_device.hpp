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

// Roe_flux_centered (flux.C:299-541) in characteristic-wave form for the
// mixed path's FP64 outer residual: the dissipation T |Lambda| Tinv dQ is
// applied as T (|Lambda| (Tinv dQ)) with the wave strengths formed directly
// and shared reciprocals (6 divisions instead of about 25, no 4x4 matrices in
// registers).  Same mathematics as roe_flux, different rounding: agrees with
// it to a few ulps of the flux magnitude, not bit for bit.
__device__ __forceinline__ void roe_flux_wave(double nx, double ny, double len, double gamma, const double* qL,
                                              const double* qR, double* flux) {
  const double g1 = gamma - 1.0;
  double irL = 1.0 / qL[0], irR = 1.0 / qR[0];
  double uL = qL[1] * irL, vL = qL[2] * irL, uR = qR[1] * irR, vR = qR[2] * irR;
  double pL = g1 * (qL[3] - 0.5 * qL[0] * (uL * uL + vL * vL));
  double pR = g1 * (qR[3] - 0.5 * qR[0] * (uR * uR + vR * vR));
  double hL = (qL[3] + pL) * irL, hR = (qR[3] + pR) * irR;
  double thL = uL * nx + vL * ny, thR = uR * nx + vR * ny;
  // Roe average
  double rb = sqrt(qL[0] * qR[0]);
  double bg = rb / (qL[0] + rb);
  double ub = uL + bg * (uR - uL), vb = vL + bg * (vR - vL), hb = hL + bg * (hR - hL);
  double th = ub * nx + vb * ny;
  double hV2 = 0.5 * (ub * ub + vb * vb);
  double c = sqrt(g1 * (hb - hV2));
  double ic = 1.0 / c, ic2 = ic * ic;
  // wave strengths Tinv dQ
  double d0 = qR[0] - qL[0], d1 = qR[1] - qL[1], d2 = qR[2] - qL[2], d3 = qR[3] - qL[3];
  double dp = g1 * (hV2 * d0 - ub * d1 - vb * d2 + d3);
  double dun = nx * d1 + ny * d2 - th * d0;
  double a0 = fabs(th) * (d0 - dp * ic2);
  double a1 = fabs(th) * ((ub * ny - vb * nx) * d0 - ny * d1 + nx * d2);
  double a2 = fabs(th + c) * 0.5 * (c * dun + dp);
  double a3 = fabs(th - c) * 0.5 * (dp - c * dun);
  // T (|Lambda| w)
  double sp = (a2 + a3) * ic2, sm = (a2 - a3) * ic;
  double D0 = a0 + sp;
  double D1 = ub * a0 - ny * a1 + ub * sp + nx * sm;
  double D2 = vb * a0 + nx * a1 + vb * sp + ny * sm;
  double D3 = hV2 * a0 + (vb * nx - ub * ny) * a1 + (1.0 / g1 + hV2 * ic2) * (a2 + a3) + th * sm;
  // physical fluxes
  double fl0 = qL[0] * thL, fr0 = qR[0] * thR;
  double fl1 = qL[1] * thL + pL * nx, fr1 = qR[1] * thR + pR * nx;
  double fl2 = qL[2] * thL + pL * ny, fr2 = qR[2] * thR + pR * ny;
  double fl3 = (qL[3] + pL) * thL, fr3 = (qR[3] + pR) * thR;
  double hl = 0.5 * len;
  flux[0] = hl * (fl0 + fr0 - D0);
  flux[1] = hl * (fl1 + fr1 - D1);
  flux[2] = hl * (fl2 + fr2 - D2);
  flux[3] = hl * (fl3 + fr3 - D3);
}

// Characteristic ghost state from the interior state and the previous ghost
// (cvbc.C:63-475).  Returns false where the reference would exit().
// The freestream (uinf, vinf, pinf) is formed on the host exactly as
// cvbc.C:62-69 does, so the device never re-evaluates cos/sin.
__device__ __forceinline__ bool ghost_state(int bc, double nx, double ny, double gamma, double uinf, double vinf,
                                            double pinf, const double* qi, double* qg) {
  const double g = gamma;
  const double rhoinf = 1.0;
  double rhoi = qi[0], ui = qi[1] / rhoi, vi = qi[2] / rhoi;
  double pi = (g - 1.) * (qi[3] - 0.5 * rhoi * (ui * ui + vi * vi));
  double ci = sqrt((g * pi) / rhoi);
  double rhob = qg[0], ub = qg[1] / rhob, vb = qg[2] / rhob;
  double pb = (g - 1.) * (qg[3] - 0.5 * rhob * (ub * ub + vb * vb));
  double cb = sqrt((g * pb) / rhob);
  double rho0 = 0.5 * (rhoi + rhob), c0 = 0.5 * (ci + cb);
  double theta = ui * nx + vi * ny;
  double ev1 = theta, ev3 = theta + ci, ev4 = theta - ci;
  if (bc == 0 || (bc % 2 == 0)) {
    if (ev1 <= 0.0 && ev4 <= 0.0) {
      if (ev3 >= 0.0) {
        pb = 0.5 * (pinf + pi - rho0 * c0 * ((uinf - ui) * nx + (vinf - vi) * ny));
        rhob = rhoinf + (pb - pinf) / (c0 * c0);
        ub = uinf + (pb - pinf) / (rho0 * c0) * nx;
        vb = vinf + (pb - pinf) / (rho0 * c0) * ny;
      } else if (ev3 < 0.0) {
        rhob = rhoinf; ub = uinf; vb = vinf; pb = pinf;
      } else {
        return false;
      }
    } else if (ev1 >= 0.0 && ev3 >= 0.0) {
      if (ev4 <= 0.0) {
        pb = pinf;
        rhob = rhoi + (pb - pi) / (c0 * c0);
        ub = ui - (pb - pi) / (rho0 * c0) * nx;
        vb = vi - (pb - pi) / (rho0 * c0) * ny;
      } else if (ev4 > 0.0) {
        rhob = rhoi; ub = ui; vb = vi; pb = pi;
      } else {
        return false;
      }
    } else {
      return false;
    }
  } else {
    pb = pi + rho0 * c0 * theta;
    rhob = rhoi + (pb - pi) / (c0 * c0);
    ub = ui - (pb - pi) / (rho0 * c0) * nx;
    vb = vi - (pb - pi) / (rho0 * c0) * ny;
  }
  if (rhob < 0.0) {
    rhob = rhoi;
    ub = (1. - 2. * nx * nx) * ui - 2. * nx * ny * vi;
    vb = (1. - 2. * ny * ny) * vi - 2. * nx * ny * ui;
    pb = pi;
  }
  qg[0] = rhob;
  qg[1] = rhob * ub;
  qg[2] = rhob * vb;
  qg[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
  return true;
}

// diagonal boundary Jacobian 0.5*(A(qL) + |A_roe|)*len (jacobian.C:681-962)
__device__ __forceinline__ void boundary_jacobian(double nx, double ny, double len, double g, const double* QL,
                                                  const double* QR, double* d) {
  double Rm[16];
  roe_matrix(nx, ny, g, QL, QR, Rm);
  double rhoL = QL[0], uL = QL[1] / QL[0], vL = QL[2] / QL[0], EL = QL[3];
  double theta = uL * nx + vL * ny;
  double phi = 0.5 * (g - 1.) * (uL * uL + vL * vL);
  d[0] = 0.; d[1] = nx; d[2] = ny; d[3] = 0.;
  d[4] = phi * nx - uL * theta;
  d[5] = (2. - g) * uL * nx + theta;
  d[6] = uL * ny - (g - 1.) * vL * nx;
  d[7] = (g - 1.) * nx;
  d[8] = phi * ny - vL * theta;
  d[9] = vL * nx - (g - 1.) * uL * ny;
  d[10] = theta + (2. - g) * vL * ny;
  d[11] = (g - 1.) * ny;
  d[12] = (2. * phi - g * EL / rhoL) * theta;
  d[13] = (1. - g) * uL * theta + nx * (g * EL / rhoL - phi);
  d[14] = (1. - g) * vL * theta + ny * (g * EL / rhoL - phi);
  d[15] = g * theta;
#pragma unroll
  for (int i = 0; i < 16; i++) d[i] = 0.5 * (d[i] + Rm[i]) * len;
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
