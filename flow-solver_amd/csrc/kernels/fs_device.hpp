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

namespace fsd {

constexpr int NV = 4;

// ---------------------------------------------------------------------------
// FP64, reference operation order
// ---------------------------------------------------------------------------

// physical flux F(q).n*len (flux.C:37-68)
__device__ __forceinline__ void euler_flux(const double* Q, double* F, double nx, double ny, double len,
                                           double gamma) {
  double rho = Q[0];
  double u = Q[1] / Q[0];
  double v = Q[2] / Q[0];
  double E = Q[3];
  double P = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
  double theta = u * nx + v * ny;
  F[0] = rho * theta * len;
  F[1] = rho * u * theta * len + P * nx * len;
  F[2] = rho * v * theta * len + P * ny * len;
  F[3] = theta * (E + P) * len;
}

// Roe matrix T |Lambda| Tinv for the Roe average of (qL, qR)
// (flux.C:356-524; MMMult_DtG / MMMult_GtG of linear_algebra.C:657-714)
__device__ __forceinline__ void roe_matrix(double nx, double ny, double gamma, const double* qleft,
                                           const double* qright, double* Roemat) {
  double rhoL = qleft[0], uL = qleft[1] / qleft[0], vL = qleft[2] / qleft[0], EL = qleft[3];
  double pL = (gamma - 1.) * (EL - 0.5 * rhoL * (uL * uL + vL * vL));
  double HL = EL + pL, hL = HL / rhoL;
  double rhoR = qright[0], uR = qright[1] / qright[0], vR = qright[2] / qright[0], ER = qright[3];
  double pR = (gamma - 1.) * (ER - 0.5 * rhoR * (uR * uR + vR * vR));
  double HR = ER + pR, hR = HR / rhoR;
  double rhobar = sqrt(rhoL * rhoR);
  double big_gamma = rhobar / (rhoL + rhobar);
  double ubar = uL + big_gamma * (uR - uL);
  double vbar = vL + big_gamma * (vR - vL);
  double hbar = hL + big_gamma * (hR - hL);
  double thetabar = ubar * nx + vbar * ny;
  double V2bar = (ubar * ubar + vbar * vbar);
  double cbar = sqrt((gamma - 1.0) * (hbar - 0.5 * V2bar));
  double Lam[4];
  Lam[0] = fabs(thetabar);
  Lam[1] = fabs(thetabar);
  Lam[2] = fabs(thetabar + cbar);
  Lam[3] = fabs(thetabar - cbar);
  double T[16], Ti[16];
  T[0] = 1.0;
  T[1] = 0.0;
  T[2] = 1.0 / (cbar * cbar);
  T[3] = 1.0 / (cbar * cbar);
  T[4] = ubar;
  T[5] = -ny;
  T[6] = ubar / (cbar * cbar) + nx / cbar;
  T[7] = ubar / (cbar * cbar) - nx / cbar;
  T[8] = vbar;
  T[9] = nx;
  T[10] = vbar / (cbar * cbar) + ny / cbar;
  T[11] = vbar / (cbar * cbar) - ny / cbar;
  T[12] = 0.5 * V2bar;
  T[13] = ubar * (-ny) + vbar * nx;
  T[14] = 1.0 / (gamma - 1.0) + thetabar / cbar + (0.5 * V2bar) / (cbar * cbar);
  T[15] = 1.0 / (gamma - 1.0) - thetabar / cbar + (0.5 * V2bar) / (cbar * cbar);
  Ti[0] = 1.0 - (gamma - 1.0) * (0.5 * V2bar) / (cbar * cbar);
  Ti[1] = (gamma - 1.0) * ubar / (cbar * cbar);
  Ti[2] = (gamma - 1.0) * vbar / (cbar * cbar);
  Ti[3] = -1.0 * (gamma - 1.0) / (cbar * cbar);
  Ti[4] = -1.0 * (ubar * (-ny) + vbar * nx);
  Ti[5] = -ny;
  Ti[6] = nx;
  Ti[7] = 0.;
  Ti[8] = 0.5 * (-cbar * thetabar + (gamma - 1.0) * 0.5 * V2bar);
  Ti[9] = 0.5 * (cbar * nx - (gamma - 1.0) * ubar);
  Ti[10] = 0.5 * (cbar * ny - (gamma - 1.0) * vbar);
  Ti[11] = 0.5 * (gamma - 1.0);
  Ti[12] = 0.5 * (cbar * thetabar + (gamma - 1.0) * 0.5 * V2bar);
  Ti[13] = -0.5 * (cbar * nx + (gamma - 1.0) * ubar);
  Ti[14] = -0.5 * (cbar * ny + (gamma - 1.0) * vbar);
  Ti[15] = 0.5 * (gamma - 1.0);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double acc = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) acc += T[i * 4 + k] * (Lam[k] * Ti[k * 4 + j]);
      Roemat[i * 4 + j] = acc;
    }
}

// Roe_flux_centered (flux.C:299-541), RECON_PRIM = 0
__device__ __forceinline__ void roe_flux(double nx, double ny, double len, double gamma, const double* qleft,
                                         const double* qright, double* flux) {
  double Rm[16];
  roe_matrix(nx, ny, gamma, qleft, qright, Rm);
  double dQ[4], diss[4], FL[4], FR[4];
#pragma unroll
  for (int i = 0; i < 4; i++) dQ[i] = qright[i] - qleft[i];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double acc = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) acc += Rm[i * 4 + j] * dQ[j];
    diss[i] = acc * len;
  }
  euler_flux(qleft, FL, nx, ny, len, gamma);
  euler_flux(qright, FR, nx, ny, len, gamma);
#pragma unroll
  for (int i = 0; i < 4; i++) flux[i] = 0.5 * (FL[i] + FR[i] - diss[i]);
}

// max(|theta+c|, |theta-c|) of one node's own state (cv_calc.C:819-849)
__device__ __forceinline__ double max_ev(const double* q, double nx, double ny, double gamma) {
  double rho = q[0], u = q[1] / rho, v = q[2] / rho;
  double P = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
  double c = sqrt((gamma * P) / rho);
  double theta = u * nx + v * ny;
  double tpc = fabs(theta + c), tmc = fabs(theta - c);
  return tpc >= tmc ? tpc : tmc;
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
