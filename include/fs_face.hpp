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

// Primitive quantities of one state as flux.C computes them (flux.C:37-46,
// flux.C:356-371): u = rho u / rho, v = rho v / rho, p, H = E + p, h = H/rho.
// Every function below reads them from here, so a state shared by several
// flux evaluations (the unperturbed side of a finite-difference Jacobian) is
// converted once with the same operations.
struct Prim {
  double rho, u, v, E, p, H, h;
};
FS_HD Prim prim(const double* q, double gamma) {
  Prim s;
  s.rho = q[0];
  s.u = q[1] / q[0];
  s.v = q[2] / q[0];
  s.E = q[3];
  s.p = (gamma - 1.) * (s.E - 0.5 * s.rho * (s.u * s.u + s.v * s.v));
  s.H = s.E + s.p;
  s.h = s.H / s.rho;
  return s;
}

// physical flux F(q).n*len (flux.C:37-68)
FS_HD void euler_flux_prim(const Prim& s, double* F, double nx, double ny, double len) {
  double theta = s.u * nx + s.v * ny;
  F[0] = s.rho * theta * len;
  F[1] = s.rho * s.u * theta * len + s.p * nx * len;
  F[2] = s.rho * s.v * theta * len + s.p * ny * len;
  F[3] = theta * (s.E + s.p) * len;
}
FS_HD void euler_flux(const double* Q, double* F, double nx, double ny, double len, double gamma) {
  euler_flux_prim(prim(Q, gamma), F, nx, ny, len);
}

// Roe matrix T |Lambda| Tinv for the Roe average of (qL, qR)
// (flux.C:356-524; MMMult_DtG / MMMult_GtG of linear_algebra.C:657-714)
FS_HD void roe_matrix_prim(double nx, double ny, double gamma, const Prim& L, const Prim& R, double* Roemat) {
  double rhobar = sqrt(L.rho * R.rho);
  double big_gamma = rhobar / (L.rho + rhobar);
  double ubar = L.u + big_gamma * (R.u - L.u);
  double vbar = L.v + big_gamma * (R.v - L.v);
  double hbar = L.h + big_gamma * (R.h - L.h);
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
FS_HD void roe_matrix(double nx, double ny, double gamma, const double* qleft, const double* qright, double* Roemat) {
  roe_matrix_prim(nx, ny, gamma, prim(qleft, gamma), prim(qright, gamma), Roemat);
}

// Roe_flux_centered (flux.C:299-541), RECON_PRIM = 0, from the two states,
// their primitive quantities and their physical fluxes
FS_HD void roe_flux_parts(double nx, double ny, double len, double gamma, const double* qleft, const double* qright,
                          const Prim& L, const Prim& R, const double* FL, const double* FR, double* flux) {
  double Rm[16];
  roe_matrix_prim(nx, ny, gamma, L, R, Rm);
  double dQ[4], diss[4];
FS_UNROLL
  for (int i = 0; i < 4; i++) dQ[i] = qright[i] - qleft[i];
FS_UNROLL
  for (int i = 0; i < 4; i++) {
    double acc = 0.;
FS_UNROLL
    for (int j = 0; j < 4; j++) acc += Rm[i * 4 + j] * dQ[j];
    diss[i] = acc * len;
  }
FS_UNROLL
  for (int i = 0; i < 4; i++) flux[i] = 0.5 * (FL[i] + FR[i] - diss[i]);
}
FS_HD void roe_flux(double nx, double ny, double len, double gamma, const double* qleft, const double* qright,
                    double* flux) {
  const Prim L = prim(qleft, gamma), R = prim(qright, gamma);
  double FL[4], FR[4];
  euler_flux_prim(L, FL, nx, ny, len);
  euler_flux_prim(R, FR, nx, ny, len);
  roe_flux_parts(nx, ny, len, gamma, qleft, qright, L, R, FL, FR, flux);
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
