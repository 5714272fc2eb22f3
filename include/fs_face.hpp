/* Per-face FP64 math of the residual path, compiled for the host and for
 * gfx950 from the same source: the Euler flux, the Roe dissipation matrix,
 * Roe_flux_centered, the node eigenvalue bound, the characteristic ghost
 * state, the flux Jacobian blocks (analytic, finite-difference, boundary)
 * and the face-state extrapolation.
 *
 * Operation order follows the reference term by term (flux.C:37-68,
 * flux.C:299-541, cv_calc.C:819-849, cvbc.C:63-475, jacobian.C:45-962,
 * reconstruct.C:181-205); compiled without floating-point
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

// Characteristic ghost state from the interior state and the previous ghost
// (cvbc.C:63-475).  Returns false where the reference would exit().
// The freestream (uinf, vinf, pinf) is formed on the host exactly as
// cvbc.C:62-69 does, so the device never re-evaluates cos/sin.
FS_HD bool ghost_state(int bc, double nx, double ny, double gamma, double uinf, double vinf,
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

// Euler flux Jacobian dF.n/dq of one state, in the operation order of
// jacobian.C:561-582 (also jacobian.C:581-599 for the right state and
// jacobian.C:892-909 at a boundary edge)
FS_HD void flux_jacobian(double nx, double ny, double g, const double* Q, double* d) {
  double rhoL = Q[0], uL = Q[1] / Q[0], vL = Q[2] / Q[0], EL = Q[3];
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
}

// diagonal boundary Jacobian 0.5*(A(qL) + |A_roe|)*len (jacobian.C:681-962)
FS_HD void boundary_jacobian(double nx, double ny, double len, double g, const double* QL, const double* QR,
                             double* d) {
  double Rm[16];
  roe_matrix(nx, ny, g, QL, QR, Rm);
  flux_jacobian(nx, ny, g, QL, d);
FS_UNROLL
  for (int i = 0; i < 16; i++) d[i] = 0.5 * (d[i] + Rm[i]) * len;
}

// compute_approximate_jacobian_roe's blocks (jacobian.C:301-612):
// dL = 0.5 (A(qL) + |A_roe|) len, dR = 0.5 (A(qR) - |A_roe|) len, row-major
FS_HD void analytic_jacobian(double nx, double ny, double len, double g, const double* QL, const double* QR,
                             double* dL, double* dR) {
  double Rm[16];
  roe_matrix(nx, ny, g, QL, QR, Rm);
  flux_jacobian(nx, ny, g, QL, dL);
  flux_jacobian(nx, ny, g, QR, dR);
FS_UNROLL
  for (int i = 0; i < 16; i++) {
    dL[i] = 0.5 * (dL[i] + Rm[i]) * len;
    dR[i] = 0.5 * (dR[i] - Rm[i]) * len;
  }
}

// compute_numerical_jacobian_roe's blocks (jacobian.C:64-106): forward
// differences of Roe_flux_centered with eps = 1e-8 in each component of qL
// (dL) and of qR (dR) around `flux` = Roe_flux_centered(qL, qR), row-major
FS_HD void fd_jacobian(double nx, double ny, double len, double g, const double* QL, const double* QR,
                       const double* flux, double* dL, double* dR) {
  const double eps = 1.0e-08;
  double Qp[4], fp[4];
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) Qp[j] = QL[j];
    Qp[i] += eps;
    roe_flux(nx, ny, len, g, Qp, QR, fp);
    for (int j = 0; j < 4; j++) dL[4 * j + i] = (fp[j] - flux[j]) / eps;
  }
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) Qp[j] = QR[j];
    Qp[i] += eps;
    roe_flux(nx, ny, len, g, QL, Qp, fp);
    for (int j = 0; j < 4; j++) dR[4 * j + i] = (fp[j] - flux[j]) / eps;
  }
}

// Face-state extrapolation of one node (reconstruct.C:181-205, 1276-1296):
// linear, Q += phi (dQ/dx dx + dQ/dy dy) with grad[var][x,y]; quadratic,
// Q = point value + phi (Q_x dx + Q_y dy) + sigma (Q_xx dx^2/2 + Q_yy dy^2/2
// + Q_xy dx dy) with hess[var * stride + (x, y, xx, yy, xy)]
FS_HD void extrapolate_linear(double* Q, double phi, const double* grad, double dx0, double dx1) {
  for (int j = 0; j < 4; j++) Q[j] += (phi * grad[j * 2 + 0] * dx0 + phi * grad[j * 2 + 1] * dx1);
}
FS_HD void extrapolate_quadratic(double* Q, const double* pv, double phi, double sigma, const double* hess, int stride,
                                 double dx0, double dx1) {
  for (int j = 0; j < 4; j++) Q[j] = pv[j];
  for (int j = 0; j < 4; j++) {
    const double* h = hess + j * stride;
    Q[j] += (phi * h[0] * dx0 + phi * h[1] * dx1 + sigma * h[2] * dx0 * dx0 * 0.5 + sigma * h[3] * dx1 * dx1 * 0.5 +
             sigma * h[4] * dx0 * dx1);
  }
}

}  // namespace fsd
