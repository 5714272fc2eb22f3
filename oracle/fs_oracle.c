/* Oracle: plain-C restatement of the reference's FP64 residual-and-implicit-
 * update path.  TEST INFRASTRUCTURE ONLY (see fs_oracle.h for who may load it
 * and how each function is pinned).  Every function cites the reference lines
 * it restates; expression order is kept so that results match the reference
 * build bit for bit where the same operations run.
 */
#include "fs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 4
#define OMIN(x, y) ((x) <= (y) ? (x) : (y))
#define OMAX(x, y) ((x) >= (y) ? (x) : (y))

/* ---- flux.C:37-68 (physical flux) ---- */
static void euler_flux(const double* Q, double* F, double nx, double ny, double len, double gamma) {
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

/* Roe matrix T*|Lambda|*Tinv from the Roe-averaged state (flux.C:356-524,
 * jacobian.C:760-860). */
static void roe_matrix(double nx, double ny, double gamma, const double* qleft, const double* qright, double* Roemat) {
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
  double Lambda[4], T[16], Tinv[16], temp[16];
  int i, j, k;
  Lambda[0] = fabs(thetabar);
  Lambda[1] = fabs(thetabar);
  Lambda[2] = fabs(thetabar + cbar);
  Lambda[3] = fabs(thetabar - cbar);
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
  Tinv[0] = 1.0 - (gamma - 1.0) * (0.5 * V2bar) / (cbar * cbar);
  Tinv[1] = (gamma - 1.0) * ubar / (cbar * cbar);
  Tinv[2] = (gamma - 1.0) * vbar / (cbar * cbar);
  Tinv[3] = -1.0 * (gamma - 1.0) / (cbar * cbar);
  Tinv[4] = -1.0 * (ubar * (-ny) + vbar * nx);
  Tinv[5] = -ny;
  Tinv[6] = nx;
  Tinv[7] = 0.;
  Tinv[8] = 0.5 * (-cbar * thetabar + (gamma - 1.0) * 0.5 * V2bar);
  Tinv[9] = 0.5 * (cbar * nx - (gamma - 1.0) * ubar);
  Tinv[10] = 0.5 * (cbar * ny - (gamma - 1.0) * vbar);
  Tinv[11] = 0.5 * (gamma - 1.0);
  Tinv[12] = 0.5 * (cbar * thetabar + (gamma - 1.0) * 0.5 * V2bar);
  Tinv[13] = -0.5 * (cbar * nx + (gamma - 1.0) * ubar);
  Tinv[14] = -0.5 * (cbar * ny + (gamma - 1.0) * vbar);
  Tinv[15] = 0.5 * (gamma - 1.0);
  /* linear_algebra.C:657-736: MMMult_DtG, MMMult_GtG */
  for (i = 0; i < 4; i++)
    for (j = 0; j < 4; j++) temp[i * 4 + j] = Lambda[i] * Tinv[i * 4 + j];
  for (i = 0; i < 4; i++)
    for (j = 0; j < 4; j++) {
      Roemat[i * 4 + j] = 0.;
      for (k = 0; k < 4; k++) Roemat[i * 4 + j] += T[i * 4 + k] * temp[k * 4 + j];
    }
}

/* flux.C:299-541 Roe_flux_centered (RECON_PRIM = 0) */
int fso_roe_flux(double nx, double ny, double len, double gamma, const double qleft[4], const double qright[4],
                 double flux[4]) {
  double Roemat[16], dQ[4], diss[4], FL[4], FR[4];
  int i, j;
  roe_matrix(nx, ny, gamma, qleft, qright, Roemat);
  for (i = 0; i < NV; i++) dQ[i] = qright[i] - qleft[i];
  for (i = 0; i < 4; i++) {
    diss[i] = 0.;
    for (j = 0; j < 4; j++) diss[i] += Roemat[i * 4 + j] * dQ[j];
  }
  for (i = 0; i < NV; i++) diss[i] = diss[i] * len;
  euler_flux(qleft, FL, nx, ny, len, gamma);
  euler_flux(qright, FR, nx, ny, len, gamma);
  for (i = 0; i < NV; i++) flux[i] = 0.5 * (FL[i] + FR[i] - diss[i]);
  return 1;
}

/* ic.C:36-99 init_pv with ic.C:808-840 ic_0 (freestream), ghosts copied */
void fso_init_freestream(const fso_mesh* m, const fso_params* p, fso_state* s) {
  double angle = p->alpha * M_PI / 180.;
  double rho = 1.0, c = 1.0;
  double P = rho * c * c / p->gamma;
  double u = p->mach * cos(angle), v = p->mach * sin(angle);
  double et = P / (p->gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  int n, i;
  for (n = 0; n < m->nn; n++) {
    s->Q[n * 4 + 0] = rho;
    s->Q[n * 4 + 1] = rho * u;
    s->Q[n * 4 + 2] = rho * v;
    s->Q[n * 4 + 3] = et;
  }
  for (i = 0; i < m->nbe; i++) {
    int r = m->bedge[5 * i + 0], g = m->nn + m->bedge[5 * i + 1];
    memcpy(&s->Q[g * 4], &s->Q[r * 4], 4 * sizeof(double));
  }
  memcpy(s->nQ, s->Q, (size_t)(m->nn + m->ng) * 4 * sizeof(double));
}

/* max |theta +- c| * len for one node's own state (cv_calc.C:819-905) */
static int max_ev(const double* q, double nx, double ny, double gamma, double* out) {
  double rho = q[0], u = q[1] / rho, v = q[2] / rho;
  double P = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
  double c = sqrt((gamma * P) / rho);
  double theta = u * nx + v * ny;
  double tpc = fabs(theta + c), tmc = fabs(theta - c);
  *out = OMAX(tpc, tmc);
  return isfinite(*out);
}

/* init_memory.C:159-310 build_matrix with cv_calc.C:790-965 */
int fso_build_matrix(const fso_mesh* m, const fso_params* p, fso_state* s) {
  int i, j, k;
  double cfl, ev;
  memset(s->LHS, 0, (size_t)m->nnz * 16 * sizeof(double));
  for (i = 0; i < m->nn; i++) s->ev_con[i] = 0.;
  for (i = 0; i < m->nsub; i++) {
    int e = m->sub_edge[i];
    double nx = m->sub_n[3 * i], ny = m->sub_n[3 * i + 1], len = m->sub_n[3 * i + 2];
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    if (!max_ev(&s->nQ[L * 4], nx, ny, p->gamma, &ev)) return -1;
    s->ev_con[L] += (ev * len);
    if (!max_ev(&s->nQ[R * 4], nx, ny, p->gamma, &ev)) return -1;
    s->ev_con[R] += (ev * len);
  }
  for (i = 0; i < m->nbe; i++) {
    int L = m->bedge[5 * i];
    double nx = m->bedge_n[3 * i], ny = m->bedge_n[3 * i + 1], len = m->bedge_n[3 * i + 2];
    if (!max_ev(&s->nQ[L * 4], nx, ny, p->gamma, &ev)) return -1;
    s->ev_con[L] += (ev * len);
  }
  if (p->method == 1) {
    if (s->citer < p->ramp)
      cfl = 1. + (1. * (s->citer - 1)) / (1. * p->ramp) * (p->CFL - 1.);
    else
      cfl = p->CFL;
  } else if (p->method == 0) {
    cfl = 1.0;
  } else {
    cfl = p->CFL;
  }
  s->cfl = cfl;
  for (j = 0; j < m->nn; j++) {
    double V = m->cv_area[j];
    i = m->iau[j];
    s->dt[j] = cfl * (V / s->ev_con[j]);
    if (!(s->dt[j] > 0.) || !isfinite(s->dt[j])) return -2;
    for (k = 0; k < NV; k++) s->LHS[(size_t)i * 16 + k * 4 + k] = V / (s->dt[j]);
  }
  return 0;
}

/* gradient.C:363-403 (edge weights from gradient.C:923-1121) */
void fso_gradient(const fso_mesh* m, fso_state* s) {
  int i, j, k;
  memset(s->grad, 0, (size_t)m->nn * 8 * sizeof(double));
  for (j = 0; j < NV; j++)
    for (i = 0; i < m->ne; i++) {
      int L = m->edges[2 * i], R = m->edges[2 * i + 1];
      const double* wl = &m->edge_weight[4 * i];
      const double* wr = &m->edge_weight[4 * i + 2];
      double dQ = s->nQ[R * 4 + j] - s->nQ[L * 4 + j];
      for (k = 0; k < 2; k++) {
        s->grad[L * 8 + 2 * j + k] += wl[k] * dQ;
        s->grad[R * 8 + 2 * j + k] -= wr[k] * dQ;
      }
    }
}

/* Weighted k=2 least-squares matrix (compute_derivatives.C:859-913) and its
 * pseudo-inverse by Householder QR (A = QR, A^+ = R^-1 Q^T); the row weights
 * w_j = 1/dist_j are folded into the columns. */
int fso_build_pinv(const fso_mesh* m, fso_state* s) {
  int i;
  for (i = 0; i < m->nn; i++) {
    int o = m->st_off[i], mm = m->st_off[i + 1] - o, r, c, j, k;
    double* A = (double*)calloc((size_t)mm * 5, sizeof(double)); /* row-major mm x 5 */
    double* Qt = (double*)calloc((size_t)mm * mm, sizeof(double));
    double* w = (double*)calloc((size_t)mm, sizeof(double));
    double xi = m->x[2 * i], yi = m->x[2 * i + 1];
    const double* Mi = &m->moments[9 * i];
    if (!A || !Qt || !w || mm < 5) { free(A); free(Qt); free(w); return -1; }
    for (r = 0; r < mm; r++) {
      int node = m->st_nodes[o + r];
      double xn = m->x[2 * node], yn = m->x[2 * node + 1];
      double dx0 = xn - xi, dx1 = yn - yi;
      double dist = sqrt(dx0 * dx0 + dx1 * dx1);
      const double* Mn = &m->moments[9 * node];
      w[r] = 1.0 / dist;
      A[r * 5 + 0] = w[r] * (xn - xi + Mn[0] - Mi[0]);
      A[r * 5 + 1] = w[r] * (yn - yi + Mn[1] - Mi[1]);
      A[r * 5 + 2] = 0.5 * w[r] * ((xn - xi) * (xn - xi) + 2.0 * Mn[0] * (xn - xi) + Mn[2] - Mi[2]);
      A[r * 5 + 3] = 0.5 * w[r] * ((yn - yi) * (yn - yi) + 2.0 * Mn[1] * (yn - yi) + Mn[3] - Mi[3]);
      A[r * 5 + 4] = w[r] * (Mn[4] + Mn[0] * (yn - yi) + Mn[1] * (xn - xi) + (yn - yi) * (xn - xi) - Mi[4]);
    }
    /* Qt starts as identity; apply Householder reflections to A and Qt */
    for (r = 0; r < mm; r++) Qt[r * mm + r] = 1.0;
    for (c = 0; c < 5; c++) {
      double norm = 0, alpha, vnorm2 = 0;
      double* v = (double*)calloc((size_t)mm, sizeof(double));
      for (r = c; r < mm; r++) norm += A[r * 5 + c] * A[r * 5 + c];
      norm = sqrt(norm);
      alpha = A[c * 5 + c] > 0 ? -norm : norm;
      for (r = c; r < mm; r++) v[r] = A[r * 5 + c];
      v[c] -= alpha;
      for (r = c; r < mm; r++) vnorm2 += v[r] * v[r];
      if (vnorm2 > 0) {
        for (k = c; k < 5; k++) {
          double d = 0;
          for (r = c; r < mm; r++) d += v[r] * A[r * 5 + k];
          d = 2.0 * d / vnorm2;
          for (r = c; r < mm; r++) A[r * 5 + k] -= d * v[r];
        }
        for (k = 0; k < mm; k++) {
          double d = 0;
          for (r = c; r < mm; r++) d += v[r] * Qt[r * mm + k];
          d = 2.0 * d / vnorm2;
          for (r = c; r < mm; r++) Qt[r * mm + k] -= d * v[r];
        }
      }
      free(v);
    }
    /* W = R^-1 (Qt)[0:5,:], back substitution column by column */
    for (j = 0; j < mm; j++) {
      double y[5];
      for (r = 4; r >= 0; r--) {
        double acc = Qt[r * mm + j];
        for (k = r + 1; k < 5; k++) acc -= A[r * 5 + k] * y[k];
        if (A[r * 5 + r] == 0.0) { free(A); free(Qt); free(w); return -2; }
        y[r] = acc / A[r * 5 + r];
      }
      for (r = 0; r < 5; r++) s->W[5 * (size_t)o + (size_t)r * mm + j] = y[r] * w[j];
    }
    free(A);
    free(Qt);
    free(w);
  }
  return 0;
}

/* compute_derivatives.C:1438-1521 (order 3 application); sol = A^+ b with
 * b_j = w_j (q_j - q_i), point value with the moment correction */
void fso_derivatives(const fso_mesh* m, fso_state* s) {
  int i, var, r, j;
  memset(s->hess, 0, (size_t)m->nn * 20 * sizeof(double));
  memset(s->pv, 0, (size_t)m->nn * 4 * sizeof(double));
  for (i = 0; i < m->nn; i++) {
    int o = m->st_off[i], mm = m->st_off[i + 1] - o;
    const double* Wi = &s->W[5 * (size_t)o];
    const double* Mi = &m->moments[9 * i];
    for (var = 0; var < NV; var++) {
      double sol[5];
      for (r = 0; r < 5; r++) {
        double acc = 0.0;
        for (j = 0; j < mm; j++) {
          int node = m->st_nodes[o + j];
          acc += Wi[r * mm + j] * (s->nQ[node * 4 + var] - s->nQ[i * 4 + var]);
        }
        sol[r] = acc;
      }
      for (r = 0; r < 5; r++) s->hess[i * 20 + var * 5 + r] = sol[r];
      s->pv[i * 4 + var] = s->nQ[i * 4 + var] - Mi[0] * sol[0] - Mi[1] * sol[1] - Mi[2] * sol[2] * 0.5 -
                           Mi[3] * sol[3] * 0.5 - Mi[4] * sol[4];
    }
  }
}

/* Reconstruction phase selection (reconstruct.C:72-171) for orders 1 and 3:
 * 0 = nodal values, 1 = linear, 3 = quadratic (k=2 phase) */
static int recon_order(const fso_params* p, int ts) {
  if (p->order == 0) return 0;
  if (p->order == 1) return ts <= p->foits ? 0 : 1;
  if (p->order == 3) {
    if (ts <= p->foits) return 0;
    if (ts <= p->foits + p->soits) return 1;
    return 3;
  }
  return -1;
}

/* unlimited quadratic reconstruction at a point (reconstruct.C:1420-1689) */
static void recon_quad_ul(const fso_state* s, const double* x, int node, const double* gp, double* q) {
  double dx0 = gp[0] - x[2 * node], dx1 = gp[1] - x[2 * node + 1];
  int j;
  for (j = 0; j < NV; j++) {
    const double* h = &s->hess[node * 20 + j * 5];
    q[j] = s->pv[node * 4 + j];
    q[j] += (h[0] * dx0 + h[1] * dx1 + h[2] * dx0 * dx0 * 0.5 + h[3] * dx1 * dx1 * 0.5 + h[4] * dx0 * dx1);
  }
}

/* one candidate update for one side: Barth min(1, r) (limiter.C:154-246) or
 * Venkatakrishnan's (r^2 + 2r) / (r^2 + r + 2) (limiter.C:696-730), with
 * r = (Qmax - q) / (qf - q) above and (Qmin - q) / (qf - q) below */
static void limiter_side(const fso_state* s, const double* qmin, const double* qmax, int node, const double* qf,
                         double* phi, int venkat) {
  int j;
  for (j = 0; j < NV; j++) {
    double q = s->nQ[node * 4 + j], d = qf[j] - q, temp;
    if (d > 0.) {
      temp = (qmax[node * 4 + j] - q) / (qf[j] - q);
      temp = venkat ? (temp * temp + 2. * temp) / (temp * temp + temp + 2.) : OMIN(1.0, temp);
      phi[node * 8 + j * 2] = OMIN(phi[node * 8 + j * 2], temp);
    } else if (d < 0.) {
      temp = (qmin[node * 4 + j] - q) / (qf[j] - q);
      temp = venkat ? (temp * temp + 2. * temp) / (temp * temp + temp + 2.) : OMIN(1.0, temp);
      phi[node * 8 + j * 2] = OMIN(phi[node * 8 + j * 2], temp);
    } else {
      phi[node * 8 + j * 2] = OMIN(phi[node * 8 + j * 2], 1.0);
    }
  }
}

/* limiter.C:39-497 compute_limiter_barth (venkat = 0) and limiter.C:511-983
 * compute_limiter_venkatakrishnan (venkat = 1): same stencil, same points */
static void limiter(const fso_mesh* m, const fso_params* p, fso_state* s, int venkat) {
  int i, j, is;
  double* qmin = (double*)malloc((size_t)m->nn * 4 * sizeof(double));
  double* qmax = (double*)malloc((size_t)m->nn * 4 * sizeof(double));
  double* phi = s->phi;
  int gauss = (p->order == 3 && s->citer > (p->foits + p->soits));
  for (i = 0; i < m->nn; i++) {
    for (j = 0; j < 8; j++) phi[i * 8 + j] = 0.0;
    for (j = 0; j < NV; j++) {
      qmin[i * 4 + j] = s->nQ[i * 4 + j];
      qmax[i * 4 + j] = s->nQ[i * 4 + j];
      phi[i * 8 + j * 2] = 1.0;
    }
  }
  for (i = 0; i < m->ne; i++) {
    int L = m->edges[2 * i], R = m->edges[2 * i + 1];
    for (j = 0; j < NV; j++) {
      qmin[L * 4 + j] = OMIN(qmin[L * 4 + j], s->nQ[R * 4 + j]);
      qmax[L * 4 + j] = OMAX(qmax[L * 4 + j], s->nQ[R * 4 + j]);
      qmin[R * 4 + j] = OMIN(qmin[R * 4 + j], s->nQ[L * 4 + j]);
      qmax[R * 4 + j] = OMAX(qmax[R * 4 + j], s->nQ[L * 4 + j]);
    }
  }
  for (is = 0; is < m->nsub; is++) {
    int e = m->sub_edge[is], g = m->sub_elem[is];
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double ql[4], qr[4];
    if (gauss) {
      int k;
      for (k = 0; k < 3; k++) {
        const double* gp = &m->sub_xg[6 * is + 2 * k];
        recon_quad_ul(s, m->x, L, gp, ql);
        recon_quad_ul(s, m->x, R, gp, qr);
        limiter_side(s, qmin, qmax, L, ql, phi, venkat);
        limiter_side(s, qmin, qmax, R, qr, phi, venkat);
      }
    } else {
      /* Reconstruct_UL, linear (reconstruct.C:600-616) */
      int ro = recon_order(p, s->citer);
      double fc0 = (m->el_cent[2 * g] + m->sub_xm[2 * is]) * 0.5;
      double fc1 = (m->el_cent[2 * g + 1] + m->sub_xm[2 * is + 1]) * 0.5;
      for (j = 0; j < NV; j++) { ql[j] = s->nQ[L * 4 + j]; qr[j] = s->nQ[R * 4 + j]; }
      if (ro == 1) {
        double dxl0 = fc0 - m->x[2 * L], dxl1 = fc1 - m->x[2 * L + 1];
        double dxr0 = fc0 - m->x[2 * R], dxr1 = fc1 - m->x[2 * R + 1];
        for (j = 0; j < NV; j++) {
          ql[j] += (s->grad[L * 8 + 2 * j] * dxl0 + s->grad[L * 8 + 2 * j + 1] * dxl1);
          qr[j] += (s->grad[R * 8 + 2 * j] * dxr0 + s->grad[R * 8 + 2 * j + 1] * dxr1);
        }
      }
      limiter_side(s, qmin, qmax, L, ql, phi, venkat);
      limiter_side(s, qmin, qmax, R, qr, phi, venkat);
    }
  }
  for (i = 0; i < m->nn; i++)
    for (j = 0; j < NV; j++) phi[i * 8 + j * 2 + 1] = phi[i * 8 + j * 2];
  for (i = 0; i < m->nn; i++) {
    double t = phi[i * 8];
    for (j = 1; j < NV; j++) t = OMIN(t, phi[i * 8 + j * 2]);
    phi[i * 8] = t;
    phi[i * 8 + 1] = t;
  }
  free(qmin);
  free(qmax);
}

void fso_limiter_barth(const fso_mesh* m, const fso_params* p, fso_state* s) { limiter(m, p, s, 0); }
void fso_limiter_venkat(const fso_mesh* m, const fso_params* p, fso_state* s) { limiter(m, p, s, 1); }

/* jacobian.C:45-185 forward-difference Jacobian around the face states */
static void numerical_jacobian(const fso_mesh* m, const fso_params* p, fso_state* s, int L, int R, const double* QL,
                               const double* QR, const double* flux, double nx, double ny, double len) {
  double Qp[4], fluxp[4], dfdqL[4][4], dfdqR[4][4];
  const double eps = 1.0e-08;
  int i, j, k, b;
  double* LHS = s->LHS;
  for (i = 0; i < NV; i++) {
    for (j = 0; j < NV; j++) Qp[j] = QL[j];
    Qp[i] += eps;
    fso_roe_flux(nx, ny, len, p->gamma, Qp, QR, fluxp);
    for (j = 0; j < NV; j++) dfdqL[j][i] = (fluxp[j] - flux[j]) / eps;
  }
  for (i = 0; i < NV; i++) {
    for (j = 0; j < NV; j++) Qp[j] = QR[j];
    Qp[i] += eps;
    fso_roe_flux(nx, ny, len, p->gamma, QL, Qp, fluxp);
    for (j = 0; j < NV; j++) dfdqR[j][i] = (fluxp[j] - flux[j]) / eps;
  }
  for (i = 0; i < NV; i++)
    for (k = 0; k < NV; k++) LHS[(size_t)m->iau[L] * 16 + i * 4 + k] += dfdqL[i][k];
  for (i = 0; i < NV; i++)
    for (k = 0; k < NV; k++) LHS[(size_t)m->iau[R] * 16 + i * 4 + k] -= dfdqR[i][k];
  for (i = m->ia[L]; i < m->ia[L + 1]; i++)
    if (m->ja[i] == R)
      for (k = 0; k < NV; k++)
        for (b = 0; b < NV; b++) LHS[(size_t)i * 16 + k * 4 + b] += dfdqR[k][b];
  for (i = m->ia[R]; i < m->ia[R + 1]; i++)
    if (m->ja[i] == L)
      for (k = 0; k < NV; k++)
        for (b = 0; b < NV; b++) LHS[(size_t)i * 16 + k * 4 + b] -= dfdqL[k][b];
}

/* jacobian.C:301-658 compute_approximate_jacobian_roe: dF/dqL = 0.5(A(qL) +
 * |A_roe|) len, dF/dqR = 0.5(A(qR) - |A_roe|) len, scattered like the
 * numerical Jacobian */
static void flux_jac_rows(double g, const double* Q, double nx, double ny, double d[4][4]) {
  double rho = Q[0], u = Q[1] / Q[0], v = Q[2] / Q[0], E = Q[3];
  double theta = u * nx + v * ny;
  double phi = 0.5 * (g - 1.) * (u * u + v * v);
  d[0][0] = 0.; d[0][1] = nx; d[0][2] = ny; d[0][3] = 0.;
  d[1][0] = phi * nx - u * theta;
  d[1][1] = (2. - g) * u * nx + theta;
  d[1][2] = u * ny - (g - 1.) * v * nx;
  d[1][3] = (g - 1.) * nx;
  d[2][0] = phi * ny - v * theta;
  d[2][1] = v * nx - (g - 1.) * u * ny;
  d[2][2] = theta + (2. - g) * v * ny;
  d[2][3] = (g - 1.) * ny;
  d[3][0] = (2. * phi - g * E / rho) * theta;
  d[3][1] = (1. - g) * u * theta + nx * (g * E / rho - phi);
  d[3][2] = (1. - g) * v * theta + ny * (g * E / rho - phi);
  d[3][3] = g * theta;
}

static void approximate_jacobian(const fso_mesh* m, const fso_params* p, fso_state* s, int L, int R, const double* QL,
                                 const double* QR, double nx, double ny, double len) {
  double Roemat[16], dL[4][4], dR[4][4];
  int i, k, b;
  double* LHS = s->LHS;
  roe_matrix(nx, ny, p->gamma, QL, QR, Roemat);
  flux_jac_rows(p->gamma, QL, nx, ny, dL);
  flux_jac_rows(p->gamma, QR, nx, ny, dR);
  for (i = 0; i < 4; i++)
    for (k = 0; k < 4; k++) {
      dL[i][k] = 0.5 * (dL[i][k] + Roemat[i * 4 + k]) * len;
      dR[i][k] = 0.5 * (dR[i][k] - Roemat[i * 4 + k]) * len;
    }
  for (i = 0; i < 4; i++)
    for (k = 0; k < 4; k++) LHS[(size_t)m->iau[L] * 16 + i * 4 + k] += dL[i][k];
  for (i = 0; i < 4; i++)
    for (k = 0; k < 4; k++) LHS[(size_t)m->iau[R] * 16 + i * 4 + k] -= dR[i][k];
  for (i = m->ia[L]; i < m->ia[L + 1]; i++)
    if (m->ja[i] == R)
      for (k = 0; k < 4; k++)
        for (b = 0; b < 4; b++) LHS[(size_t)i * 16 + k * 4 + b] += dR[k][b];
  for (i = m->ia[R]; i < m->ia[R + 1]; i++)
    if (m->ja[i] == L)
      for (k = 0; k < 4; k++)
        for (b = 0; b < 4; b++) LHS[(size_t)i * 16 + k * 4 + b] -= dL[k][b];
}

static int g_analytic_jacobian = 0;
void fso_set_analytic_jacobian(int on) { g_analytic_jacobian = on; }

/* residual.C:40-342 build_interior_residuals */
void fso_interior_residual(const fso_mesh* m, const fso_params* p, fso_state* s, int jupdate, int ts) {
  int is, j;
  int ro = recon_order(p, ts);
  for (j = 0; j < m->nn * 4; j++) { s->RHS[j] = 0.; s->R[j] = 0.; }
  for (is = 0; is < m->nsub; is++) {
    int e = m->sub_edge[is], g = m->sub_elem[is];
    double nx = m->sub_n[3 * is], ny = m->sub_n[3 * is + 1], len = m->sub_n[3 * is + 2];
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double ql[4], qr[4], flux[4];
    for (j = 0; j < NV; j++) { ql[j] = s->nQ[L * 4 + j]; qr[j] = s->nQ[R * 4 + j]; }
    if (ro == 1) {
      /* reconstruct.C:181-205 (limited, linear) */
      double fc0 = (m->el_cent[2 * g] + m->sub_xm[2 * is]) * 0.5;
      double fc1 = (m->el_cent[2 * g + 1] + m->sub_xm[2 * is + 1]) * 0.5;
      double dxl0 = fc0 - m->x[2 * L], dxl1 = fc1 - m->x[2 * L + 1];
      double dxr0 = fc0 - m->x[2 * R], dxr1 = fc1 - m->x[2 * R + 1];
      double phiL = s->phi[L * 8], phiR = s->phi[R * 8];
      for (j = 0; j < NV; j++) {
        ql[j] += (phiL * s->grad[L * 8 + 2 * j] * dxl0 + phiL * s->grad[L * 8 + 2 * j + 1] * dxl1);
        qr[j] += (phiR * s->grad[R * 8 + 2 * j] * dxr0 + phiR * s->grad[R * 8 + 2 * j + 1] * dxr1);
      }
    }
    /* ro == 3: the Gauss-point flux is overwritten by the nodal flux
     * (residual.C:208-290), so only the nodal flux is formed here */
    fso_roe_flux(nx, ny, len, p->gamma, ql, qr, flux);
    if (jupdate && g_analytic_jacobian) approximate_jacobian(m, p, s, L, R, ql, qr, nx, ny, len);
    else if (jupdate) numerical_jacobian(m, p, s, L, R, ql, qr, flux, nx, ny, len);
    for (j = 0; j < NV; j++) {
      s->R[L * 4 + j] += flux[j];
      s->R[R * 4 + j] -= flux[j];
    }
  }
}

/* cvbc.C:37-481 Set_Boundary_Values (farfield and inviscid wall) */
static int set_boundary_values(const fso_mesh* m, const fso_params* p, fso_state* s) {
  double angle = p->alpha * M_PI / 180.;
  double rhoinf = 1.0, uinf = p->mach * cos(angle), vinf = p->mach * sin(angle), pinf = 1. / p->gamma;
  double g = p->gamma;
  int i;
  for (i = 0; i < m->nbe; i++) {
    int real = m->bedge[5 * i], gh = m->nn + m->bedge[5 * i + 1], b = m->bedge[5 * i + 3];
    int bc = m->bbc[b];
    double nx = m->bedge_n[3 * i], ny = m->bedge_n[3 * i + 1];
    double* qi = &s->nQ[real * 4];
    double* qg = &s->nQ[gh * 4];
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
          return -1;
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
          return -1;
        }
      } else {
        return -1;
      }
    } else {
      pb = pi + rho0 * c0 * theta;
      rhob = rhoi + (pb - pi) / (c0 * c0);
      ub = ui - (pb - pi) / (rho0 * c0) * nx;
      vb = vi - (pb - pi) / (rho0 * c0) * ny;
    }
    qg[0] = rhob;
    qg[1] = rhob * ub;
    qg[2] = rhob * vb;
    qg[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
    if (rhob < 0.0) {
      rhob = rhoi;
      ub = (1. - 2. * nx * nx) * ui - 2. * nx * ny * vi;
      vb = (1. - 2. * ny * ny) * vi - 2. * nx * ny * ui;
      pb = pi;
      qg[0] = rhob;
      qg[1] = rhob * ub;
      qg[2] = rhob * vb;
      qg[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
    }
  }
  return 0;
}

/* jacobian.C:681-962: diagonal block 0.5*(A(q_L) + |A_roe|)*len */
static void boundary_jacobian(const fso_mesh* m, const fso_params* p, fso_state* s, int ib, const double* QL,
                              const double* QR, double nx, double ny, double len) {
  double Roemat[16], d[4][4];
  double g = p->gamma;
  double rhoL = QL[0], uL = QL[1] / QL[0], vL = QL[2] / QL[0], EL = QL[3];
  double theta, phi;
  int i, k, n;
  roe_matrix(nx, ny, g, QL, QR, Roemat);
  theta = uL * nx + vL * ny;
  phi = 0.5 * (g - 1.) * (uL * uL + vL * vL);
  d[0][0] = 0.; d[0][1] = nx; d[0][2] = ny; d[0][3] = 0.;
  d[1][0] = phi * nx - uL * theta;
  d[1][1] = (2. - g) * uL * nx + theta;
  d[1][2] = uL * ny - (g - 1.) * vL * nx;
  d[1][3] = (g - 1.) * nx;
  d[2][0] = phi * ny - vL * theta;
  d[2][1] = vL * nx - (g - 1.) * uL * ny;
  d[2][2] = theta + (2. - g) * vL * ny;
  d[2][3] = (g - 1.) * ny;
  d[3][0] = (2. * phi - g * EL / rhoL) * theta;
  d[3][1] = (1. - g) * uL * theta + nx * (g * EL / rhoL - phi);
  d[3][2] = (1. - g) * vL * theta + ny * (g * EL / rhoL - phi);
  d[3][3] = g * theta;
  for (i = 0; i < 4; i++)
    for (k = 0; k < 4; k++) d[i][k] = 0.5 * (d[i][k] + Roemat[i * 4 + k]) * len;
  n = m->bedge[5 * ib];
  for (i = 0; i < 4; i++)
    for (k = 0; k < 4; k++) s->LHS[(size_t)m->iau[n] * 16 + i * 4 + k] += d[i][k];
}

/* residual.C:359-1230 build_boundary_residuals (linear boundaries: the
 * residual uses the nodal real and ghost states in every phase) */
int fso_boundary_residual(const fso_mesh* m, const fso_params* p, fso_state* s, int jupdate) {
  int i, j;
  if (set_boundary_values(m, p, s) != 0) return -1;
  for (i = 0; i < m->nbe; i++) {
    int real = m->bedge[5 * i], gh = m->nn + m->bedge[5 * i + 1];
    double nx = m->bedge_n[3 * i], ny = m->bedge_n[3 * i + 1], len = m->bedge_n[3 * i + 2];
    double ql[4], qr[4], flux[4];
    for (j = 0; j < NV; j++) { ql[j] = s->nQ[real * 4 + j]; qr[j] = s->nQ[gh * 4 + j]; }
    fso_roe_flux(nx, ny, len, p->gamma, ql, qr, flux);
    if (jupdate) boundary_jacobian(m, p, s, i, ql, qr, nx, ny, len);
    for (j = 0; j < NV; j++) s->R[real * 4 + j] += flux[j];
  }
  return 0;
}

/* linear_algebra.C:571-638 lu / forwardstab / backstab */
static void lu4(double* A) {
  int i, j, k;
  for (k = 0; k < 4; k++) {
    for (i = k + 1; i < 4; i++) A[i * 4 + k] = A[i * 4 + k] / A[k * 4 + k];
    for (j = k + 1; j < 4; j++)
      for (i = k + 1; i < 4; i++) A[i * 4 + j] = A[i * 4 + j] - A[i * 4 + k] * A[k * 4 + j];
  }
}
static void lu_solve4(const double* A, const double* b, double* x) {
  double y[4];
  y[0] = b[0];
  y[1] = b[1] - A[4] * y[0];
  y[2] = b[2] - A[8] * y[0] - A[9] * y[1];
  y[3] = b[3] - A[12] * y[0] - A[13] * y[1] - A[14] * y[2];
  x[3] = y[3] / A[15];
  x[2] = (y[2] - A[11] * x[3]) / A[10];
  x[1] = (y[1] - A[6] * x[2] - A[7] * x[3]) / A[5];
  x[0] = (y[0] - A[1] * x[1] - A[2] * x[2] - A[3] * x[3]) / A[0];
}

static void gs_row(const fso_mesh* m, fso_state* s, int j) {
  double rhs1[4], M[16], X[4], dX[4];
  int k, v, d = m->iau[j];
  const double* LHS = s->LHS;
  double* dQ = s->dQ;
  for (v = 0; v < 4; v++) rhs1[v] = s->RHS[j * 4 + v];
  for (k = m->ia[j]; k < m->ia[j + 1]; k++) {
    const double* A = &LHS[(size_t)k * 16];
    const double* q = &dQ[m->ja[k] * 4];
    if (m->ja[k] == j) continue;
    for (v = 0; v < 4; v++) rhs1[v] -= (A[v * 4 + 0] * q[0] + A[v * 4 + 1] * q[1] + A[v * 4 + 2] * q[2] + A[v * 4 + 3] * q[3]);
  }
  memcpy(M, &LHS[(size_t)d * 16], 16 * sizeof(double));
  lu4(M);
  lu_solve4(M, rhs1, X);
  for (v = 0; v < 4; v++) dX[v] = X[v] - dQ[j * 4 + v];
  for (v = 0; v < 4; v++) dQ[j * 4 + v] += dX[v];
}

/* linear_algebra.C:122-401 linear_solve (symmetric block Gauss-Seidel) */
double fso_linear_solve(const fso_mesh* m, const fso_params* p, fso_state* s, int multicolor) {
  int i, j, k, c;
  double RMS = 10.;
  for (i = 0; i < p->subits; i++) {
    RMS = 0.;
    if (!multicolor) {
      if (i % 2 == 0)
        for (j = 0; j < m->nn; j++) gs_row(m, s, j);
      else
        for (j = m->nn - 1; j >= 0; j--) gs_row(m, s, j);
    } else {
      /* rows of one color are independent: visiting them in any order gives
       * the same result, so this is the device sweep's sequence */
      if (i % 2 == 0) {
        for (c = 0; c < m->ncolor; c++)
          for (k = m->color_off[c]; k < m->color_off[c + 1]; k++) gs_row(m, s, m->color_rows[k]);
      } else {
        for (c = m->ncolor - 1; c >= 0; c--)
          for (k = m->color_off[c]; k < m->color_off[c + 1]; k++) gs_row(m, s, m->color_rows[k]);
      }
    }
    for (j = 0; j < m->nn; j++) {
      double r[4];
      int v;
      for (v = 0; v < 4; v++) r[v] = s->RHS[j * 4 + v];
      for (k = m->ia[j]; k < m->ia[j + 1]; k++) {
        const double* A = &s->LHS[(size_t)k * 16];
        const double* q = &s->dQ[m->ja[k] * 4];
        for (v = 0; v < 4; v++) r[v] -= (A[v * 4 + 0] * q[0] + A[v * 4 + 1] * q[1] + A[v * 4 + 2] * q[2] + A[v * 4 + 3] * q[3]);
      }
      RMS += (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
    }
    RMS /= (1. * m->nn);
    RMS = sqrt(RMS);
    if (RMS < 1.0E-15) break;
  }
  s->last_lin = RMS;
  return RMS;
}

/* solve.C:44-933 one pseudo-time step (steady, implicit) */
double fso_solve(const fso_mesh* m, const fso_params* p, fso_state* s, int ts, int multicolor) {
  int i, j, it;
  double dql2 = 0.;
  int ntot = (m->nn + m->ng) * 4;
  s->citer = ts;
  memset(s->grad, 0, (size_t)m->nn * 8 * sizeof(double));
  memset(s->hess, 0, (size_t)m->nn * 20 * sizeof(double));
  memcpy(s->nQ, s->Q, (size_t)ntot * sizeof(double));
  for (it = 1; it <= p->iter; it++) {
    int jac = ((it - 1) % p->jufreq == 0) ? 1 : 0;
    int do_lim = 0;
    double rms = 0.;
    if (it == 1) jac = 1;
    if (p->method == 0) jac = 0;
    for (i = 0; i < m->nn * 8; i++) s->phi[i] = 1.;
    if (fso_build_matrix(m, p, s) != 0) return -1.0;
    if (p->order == 1 && ts > p->foits) fso_gradient(m, s);
    else if (p->order == 3) {
      if (ts > p->foits && ts <= p->foits + p->soits) fso_gradient(m, s);
      else if (ts > p->foits + p->soits) fso_derivatives(m, s);
    }
    if (p->order == 1 && p->limit > 0 && ts > p->foits) do_lim = 1;
    if (p->order == 3 && p->limit > 0 && ts > p->foits + p->soits) do_lim = 1;
    if (do_lim && p->limit == 1) fso_limiter_barth(m, p, s);
    if (do_lim && p->limit == 2) fso_limiter_venkat(m, p, s);
    fso_interior_residual(m, p, s, jac, ts);
    if (fso_boundary_residual(m, p, s, jac) != 0) return -1.0;
    for (i = 0; i < m->nn * 4; i++) rms += s->R[i] * s->R[i];
    s->last_rms = rms;
    memset(s->dQ, 0, (size_t)m->nn * 4 * sizeof(double));
    for (i = 0; i < m->nn * 4; i++) s->RHS[i] = -(s->R[i]);
    fso_linear_solve(m, p, s, multicolor);
    dql2 = 0.;
    for (i = 0; i < m->nn; i++)
      for (j = 0; j < 4; j++) dql2 += (s->dQ[i * 4 + j] * s->dQ[i * 4 + j]);
    for (i = 0; i < m->nn; i++)
      for (j = 0; j < 4; j++) s->nQ[i * 4 + j] += s->dQ[i * 4 + j];
  }
  memcpy(s->pQ, s->Q, (size_t)ntot * sizeof(double));
  memcpy(s->Q, s->nQ, (size_t)ntot * sizeof(double));
  return sqrt(dql2);
}

/* Reconstruct_Gauss_Boundary (reconstruct.C:1708-1890) with PV_EXT = 1,
 * RECON_PRIM = 0 and node_order = p->order: the phase of step `citer` picks
 * nodal values (0), the limited linear reconstruction (1) or the k=2
 * point-value quadratic (3). */
static void recon_gauss_boundary(const fso_mesh* m, const fso_params* p, const fso_state* s, int citer, int node,
                                 const double* gp, double* q) {
  int j, rorder = 0;
  double phi = s->phi[node * 8 + 0], sigma = s->phi[node * 8 + 1];
  double dx0, dx1;
  for (j = 0; j < NV; j++) q[j] = s->nQ[node * 4 + j];
  if (p->order == 1) rorder = citer <= p->foits ? 0 : 1;
  else if (p->order == 3) {
    if (citer <= p->foits) rorder = 0;
    else if (citer <= p->foits + p->soits) rorder = 1;
    else {
      rorder = 3;
      for (j = 0; j < NV; j++) q[j] = s->pv[node * 4 + j];
    }
  }
  dx0 = gp[0] - m->x[2 * node];
  dx1 = gp[1] - m->x[2 * node + 1];
  if (rorder == 1) {
    for (j = 0; j < NV; j++)
      q[j] += (phi * s->grad[node * 8 + j * 2 + 0] * dx0 + phi * s->grad[node * 8 + j * 2 + 1] * dx1);
  } else if (rorder == 3) {
    const double* h = &s->hess[node * 20];
    for (j = 0; j < NV; j++)
      q[j] += (phi * h[j * 5 + 0] * dx0 + phi * h[j * 5 + 1] * dx1 + sigma * h[j * 5 + 2] * dx0 * dx0 * 0.5 +
               sigma * h[j * 5 + 3] * dx1 * dx1 * 0.5 + sigma * h[j * 5 + 4] * dx0 * dx1);
  }
}

static double pressure_of(double gamma, const double* q) {
  double r = q[0], u = q[1] / r, v = q[2] / r, e = q[3];
  return (gamma - 1.) * (e - 0.5 * r * (u * u + v * v));
}

void fso_body_forces(const fso_mesh* m, const fso_params* p, const fso_state* s, int citer, double forces[2]) {
  int i, g;
  const double angle = p->alpha * M_PI / 180;
  const double t[3] = {sqrt(15.0) / 5.0, 0., sqrt(15.0) / (-5.0)};
  const double w[3] = {5. / 9., 8. / 9., 5. / 9.};
  forces[0] = 0.;
  forces[1] = 0.;
  for (i = 0; i < m->nbe; i++) {
    const int* be = &m->bedge[5 * i];
    int real = be[0], b = be[3], seg = m->bseg_off[b] + be[4];
    int bc = m->bbc[b];
    int nL, nR;
    double nx, ny, len, xi, yi, xmid, ymid, xL, xR, yL, yR, temp1 = 0., temp2 = 0.;
    if (bc == 0 || (bc % 2 == 0)) continue;
    len = m->bedge_n[3 * i + 2];
    if (bc > 10) {
      /* grid_io.C:1612-1660: curved wall, the curve's Gauss points and
         normals (from the mesh's geometry), weights 8/9, 5/9, 5/9 */
      const double wc[3] = {8. / 9., 5. / 9., 5. / 9.};
      for (g = 0; g < 3; g++) {
        double q[4], P;
        const double* gp = &m->bedge_xg[6 * i + 2 * g];
        nx = -m->bedge_gn[6 * i + 2 * g];
        ny = -m->bedge_gn[6 * i + 2 * g + 1];
        recon_gauss_boundary(m, p, s, citer, real, gp, q);
        P = pressure_of(p->gamma, q);
        temp1 += (wc[g] * (P * ny * cos(angle) - P * nx * sin(angle)));
        temp2 += (wc[g] * (P * nx * cos(angle) + P * ny * sin(angle)));
      }
      temp1 = temp1 * len * 0.5;
      temp2 = temp2 * len * 0.5;
      forces[0] += temp1;
      forces[1] += temp2;
      continue;
    }
    nx = -m->bedge_n[3 * i];
    ny = -m->bedge_n[3 * i + 1];
    nL = m->bseg[2 * seg];
    nR = m->bseg[2 * seg + 1];
    xi = m->x[2 * real];
    yi = m->x[2 * real + 1];
    xmid = 0.5 * (m->x[2 * nL] + m->x[2 * nR]);
    ymid = 0.5 * (m->x[2 * nL + 1] + m->x[2 * nR + 1]);
    if (real == nL) { xL = xi; xR = xmid; yL = yi; yR = ymid; }
    else { xL = xmid; xR = xi; yL = ymid; yR = yi; }
    for (g = 0; g < 3; g++) {
      double gp[2], q[4], P;
      gp[0] = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t[g];
      gp[1] = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t[g];
      recon_gauss_boundary(m, p, s, citer, real, gp, q);
      P = pressure_of(p->gamma, q);
      temp1 += (w[g] * (P * ny * cos(angle) - P * nx * sin(angle)));
      temp2 += (w[g] * (P * nx * cos(angle) + P * ny * sin(angle)));
    }
    temp1 = temp1 * len * 0.5;
    temp2 = temp2 * len * 0.5;
    forces[0] += temp1;
    forces[1] += temp2;
  }
}

int fso_surface_cp(const fso_mesh* m, const fso_params* p, const fso_state* s, int b, double* x, double* cp) {
  int sgi, k, e, count = 0;
  const double pinf = 1. / p->gamma, minf = p->mach;
  int bc = m->bbc[b];
  int* seen;
  if (bc == 0 || (bc % 2 == 0)) return 0;
  seen = (int*)malloc(sizeof(int) * (size_t)(2 * (m->bseg_off[b + 1] - m->bseg_off[b]) + 1));
  for (sgi = m->bseg_off[b]; sgi < m->bseg_off[b + 1]; sgi++)
    for (e = 0; e < 2; e++) {
      int n = m->bseg[2 * sgi + e], unique = 1;
      for (k = 0; k < count; k++)
        if (seen[k] == n) unique = 0;
      if (unique) {
        double P = pressure_of(p->gamma, &s->Q[n * 4]);
        seen[count] = n;
        x[count] = m->x[2 * n];
        cp[count] = (P / pinf - 1.) * (2. / (p->gamma * minf * minf));
        count++;
      }
    }
  free(seen);
  return count;
}
