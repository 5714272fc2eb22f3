/* Per-face FP64 math of the residual path, compiled for the host and for
 * gfx950 from the same source: the Euler flux, the Roe dissipation matrix,
 * Roe_flux_centered and the node eigenvalue bound.
 *
 * Operation order follows the reference term by term (flux.C:37-68,
 * flux.C:299-541, cv_calc.C:819-849); compiled without floating-point
 * contraction (hipcc -ffp-contract=off; g++ on x86-64 without an FMA target)
 * a host call and a device call give the reference's result bit for bit.
 */
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FS_HD __host__ __device__ __forceinline__
#define FS_UNROLL _Pragma("unroll")
#else
#define FS_HD inline
#define FS_UNROLL
#endif

namespace fsd {

constexpr int NV = 4;

// physical flux F(q).n*len (flux.C:37-68)
FS_HD void euler_flux(const double* Q, double* F, double nx, double ny, double len, double gamma) {
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
FS_HD void roe_matrix(double nx, double ny, double gamma, const double* qleft, const double* qright,
                      double* Roemat) {
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
FS_UNROLL
  for (int i = 0; i < 4; i++)
FS_UNROLL
    for (int j = 0; j < 4; j++) {
      double acc = 0.;
FS_UNROLL
      for (int k = 0; k < 4; k++) acc += T[i * 4 + k] * (Lam[k] * Ti[k * 4 + j]);
      Roemat[i * 4 + j] = acc;
    }
}

// Roe_flux_centered (flux.C:299-541), RECON_PRIM = 0
FS_HD void roe_flux(double nx, double ny, double len, double gamma, const double* qleft, const double* qright,
                    double* flux) {
  double Rm[16];
  roe_matrix(nx, ny, gamma, qleft, qright, Rm);
  double dQ[4], diss[4], FL[4], FR[4];
FS_UNROLL
  for (int i = 0; i < 4; i++) dQ[i] = qright[i] - qleft[i];
FS_UNROLL
  for (int i = 0; i < 4; i++) {
    double acc = 0.;
FS_UNROLL
    for (int j = 0; j < 4; j++) acc += Rm[i * 4 + j] * dQ[j];
    diss[i] = acc * len;
  }
  euler_flux(qleft, FL, nx, ny, len, gamma);
  euler_flux(qright, FR, nx, ny, len, gamma);
FS_UNROLL
  for (int i = 0; i < 4; i++) flux[i] = 0.5 * (FL[i] + FR[i] - diss[i]);
}

// max(|theta+c|, |theta-c|) of one node's own state (cv_calc.C:819-849)
FS_HD double max_ev(const double* q, double nx, double ny, double gamma) {
  double rho = q[0], u = q[1] / rho, v = q[2] / rho;
  double P = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
  double c = sqrt((gamma * P) / rho);
  double theta = u * nx + v * ny;
  double tpc = fabs(theta + c), tmc = fabs(theta - c);
  return tpc >= tmc ? tpc : tmc;
}

}  // namespace fsd
