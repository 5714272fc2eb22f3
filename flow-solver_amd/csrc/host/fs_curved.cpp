// Analytic curved walls, curved_boundaries.C:28-1155.  Each function cites
// the reference function it restates; the operation order of every formula
// follows the reference's so that the circle geometry (which needs no
// quadrature) is the same bit for bit.
#include "fs_curved.hpp"

#include <cmath>

namespace fs {
namespace {

// radius of the circle codes; 0 for the NACA codes
double circle_radius(int bc) {
  switch (bc) {
    case 11: case 12: return 1.0;
    case 13: case 14: return 2.0;
    case 15: case 16: return 3.0;
    case 21: case 22: return 300.0;
    default: return 0.0;
  }
}
bool is_naca(int bc) { return bc >= 17 && bc <= 20; }

// angle of a point of the unit circle (codes 11/12): acos of x, negative
// below the x axis (curved_boundaries.C:116-132)
double unit_angle(const double* x) { return x[1] >= 0. ? std::acos(x[0]) : -std::acos(x[0]); }
// angle of a point of the other circles
double angle_of(int bc, const double* x) {
  return (bc == 11 || bc == 12) ? unit_angle(x) : std::atan2(x[1], x[0]);
}
// end angles of a segment with mixed signs are put on one side
// (curved_boundaries.C:165-168)
void same_side(double* tL, double* tR) {
  if (*tL < 0. && *tR > 0.)
    *tR = -1. * *tR;
  else if (*tL > 0. && *tR < 0.)
    *tL = -1. * *tL;
}

// slope dy/dx of the upper surface (the integrand's derivative term,
// curved_boundaries.C:58-65)
double naca_slope(double x) {
  return 0.08907 / std::sqrt(x) - 0.07560 - 0.42192 * x + 0.51174 * x * x - 0.24864 * x * x * x;
}
// -sqrt(1 + y'^2): derivative of (target length - length to x) in x
// (naca_arclen_function2, curved_boundaries.C:67-75)
double neg_arclen_density(double x) {
  double yp = naca_slope(x);
  return -1.0 * std::sqrt(1.0 + yp * yp);
}

// Gauss-Legendre rule of order kGL on [-1, 1], nodes found by Newton's
// iteration on the three-term recurrence of the Legendre polynomials.
constexpr int kGL = 10;
struct GaussLegendre {
  double x[kGL], w[kGL];
  GaussLegendre() {
    const double PI = 3.14159265358979323846;
    for (int i = 0; i < (kGL + 1) / 2; ++i) {
      double z = std::cos(PI * (i + 0.75) / (kGL + 0.5)), dp = 1.0;
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = z;
        for (int j = 2; j <= kGL; ++j) {
          double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
          p0 = p1;
          p1 = p2;
        }
        dp = kGL * (z * p1 - p0) / (z * z - 1.0);
        double dz = p1 / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-16) break;
      }
      x[i] = -z;
      x[kGL - 1 - i] = z;
      w[i] = w[kGL - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
  }
};
const GaussLegendre& gauss_legendre() {
  static const GaussLegendre g;
  return g;
}

// Arclength density after x = t^2: sqrt(1 + y'(x)^2) dx = 2 sqrt(t^2 + (t y')^2) dt,
// where t y'(t^2) = 0.08907 - 0.0756 t - 0.42192 t^3 + 0.51174 t^5 - 0.24864 t^7
// is a polynomial: the integrand is smooth on [0, 1] and positive.
double arclen_density_t(double t) {
  const double t2 = t * t;
  const double p = 0.08907 - t * (0.07560 + t2 * (0.42192 - t2 * (0.51174 - t2 * 0.24864)));
  return 2.0 * std::sqrt(t2 + p * p);
}
// the reference's integrand in x, sqrt(1 + y'(x)^2) (curved_boundaries.C:58-65)
double arclen_density_x(double x) {
  const double yp = naca_slope(x);
  return std::sqrt(1.0 + yp * yp);
}
template <double (*F)(double)>
double panel(double a, double b) {
  const GaussLegendre& g = gauss_legendre();
  const double m = 0.5 * (a + b), h = 0.5 * (b - a);
  double s = 0.0;
  for (int i = 0; i < kGL; ++i) s += g.w[i] * F(m + h * g.x[i]);
  return s * h;
}
// Interval halving until the two halves agree with the whole to 1e-14 of
// the panel or to its share of 1e-15 of the whole integral.  The integrand
// is positive, so panel sums never cancel, and the 1e-14 floor lies well
// above the rounding of a 10-point sum (about 1e-15), so rounding alone never
// drives the recursion; the depth cap bounds it regardless.
template <double (*F)(double)>
double adapt(double a, double b, double whole, double tol, int depth) {
  const double m = 0.5 * (a + b);
  const double l = panel<F>(a, m), r = panel<F>(m, b), lr = l + r;
  const double diff = std::fabs(lr - whole);
  if (depth >= 24 || diff <= tol || diff <= 1e-14 * lr) return lr;
  return adapt<F>(a, m, l, 0.5 * tol, depth + 1) + adapt<F>(m, b, r, 0.5 * tol, depth + 1);
}
template <double (*F)(double)>
double integrate(double a, double b) {
  const double whole = panel<F>(a, b);
  return adapt<F>(a, b, whole, 1e-15 * whole, 0);
}

}  // namespace

double naca_y(double x, int bc) {
  double y = 0.6 * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
  if (bc == 19 || bc == 20) y = -1.0 * y;
  return y;
}

double naca_arclength(double xa, double xb) {
  if (!(xb > xa)) return 0.0;
  // Away from the leading edge (the interval at least twice its length from
  // x = 0) the integrand in x is smooth on the interval and is integrated as
  // it is: there sqrt(xb) - sqrt(xa) would lose digits to cancellation.
  // Otherwise in t = sqrt(x), where the singularity is gone.
  if (xa >= 2.0 * (xb - xa)) return integrate<arclen_density_x>(xa, xb);
  return integrate<arclen_density_t>(std::sqrt(xa), std::sqrt(xb));
}

void curved_midpoint(int bc, const double* xL, const double* xR, double* xM) {
  if (is_naca(bc)) {
    // curved_boundaries.C:213-218: the profile point above the chord midpoint
    xM[0] = 0.5 * (xL[0] + xR[0]);
    xM[1] = naca_y(xM[0], bc);
    return;
  }
  const double R = circle_radius(bc);
  if (R == 0.0) return;
  double tL = angle_of(bc, xL), tR = angle_of(bc, xR);
  same_side(&tL, &tR);
  const double tM = 0.5 * (tL + tR);
  if (bc == 11 || bc == 12) {
    xM[0] = std::cos(tM);
    xM[1] = std::sin(tM);
  } else {
    xM[0] = R * std::cos(tM);
    xM[1] = R * std::sin(tM);
  }
}

bool curved_gauss_points(int bc, const double* xL, const double* xR, double* GP, std::string* err) {
  const double t[3] = {0.0, std::sqrt(0.6), -1.0 * std::sqrt(0.6)};
  if (is_naca(bc)) {
    // curved_boundaries.C:550-628: the points at arclength fractions beta
    // from the end of smaller x, by Newton's iteration on x
    const double beta[3] = {0.5, (1.0 + std::sqrt(3. / 5.)) * 0.5, (1.0 - std::sqrt(3. / 5.)) * 0.5};
    double xa, xb;
    if (xL[0] < xR[0]) { xa = xL[0]; xb = xR[0]; }
    else { xa = xR[0]; xb = xL[0]; }
    const double Lt = naca_arclength(xa, xb);
    for (int i = 0; i < 3; ++i) {
      double xn = (xa + xb) * 0.5 + (xb - xa) * 0.5 * t[i], xnp1 = xn, f = 0.0;
      int j = 1;
      for (; j <= 50; ++j) {
        f = (beta[i] * Lt) - naca_arclength(xa, xn);
        const double fp = neg_arclen_density(xn);
        double update = (f / fp);
        xnp1 = xn - update;
        // the step is halved until it stays on the segment; it reaches 0
        for (int h = 0; h < 2100 && (xnp1 < xa || xnp1 > xb); ++h) {
          update = 0.5 * update;
          xnp1 = xn - update;
        }
        if (std::fabs(f) < 1.0E-12) break;
        xn = xnp1;
      }
      if (j > 50) {
        if (err) *err = "curved_gauss_points: Newton's iteration for a NACA 0012 Gauss point did not converge";
        return false;
      }
      GP[2 * i] = xnp1;
      GP[2 * i + 1] = naca_y(xnp1, bc);
    }
    return true;
  }
  const double R = circle_radius(bc);
  if (R == 0.0) {
    if (err) *err = "curved_gauss_points: unknown curved boundary code";
    return false;
  }
  double thL = angle_of(bc, xL), thR = angle_of(bc, xR);
  same_side(&thL, &thR);
  for (int k = 0; k < 3; ++k) {
    const double th = (thL + thR) * 0.5 + (thR - thL) * 0.5 * t[k];
    if (bc == 11 || bc == 12) {
      GP[2 * k] = std::cos(th);
      GP[2 * k + 1] = std::sin(th);
    } else {
      GP[2 * k] = R * std::cos(th);
      GP[2 * k + 1] = R * std::sin(th);
    }
  }
  return true;
}

void curved_normal(int bc, const double* x, double* nx, double* ny) {
  if (bc == 17 || bc == 18 || bc == 19 || bc == 20) {
    // curved_boundaries.C:787-840: normal of the profile's tangent, turned
    // into the airfoil
    if (std::fabs(x[0]) < 1.e-10) {
      *nx = 1.0;
      *ny = 0.0;
      return;
    }
    const bool upper = bc <= 18;
    double y;
    if (upper)
      y = 0.08907 / std::sqrt(x[0]) - 0.07560 - 0.42192 * x[0] + 0.51174 * x[0] * x[0] - 0.24864 * x[0] * x[0] * x[0];
    else
      y = -0.08907 / std::sqrt(x[0]) + 0.07560 + 0.42192 * x[0] - 0.51174 * x[0] * x[0] + 0.24864 * x[0] * x[0] * x[0];
    const double m = -1.0 / y;
    const double len = std::sqrt(1.0 + m * m);
    *nx = 1.0 / len;
    *ny = m / len;
    if (upper ? (*ny > 0.0) : (*ny < 0.0)) {
      *nx = -1.0 * (*nx);
      *ny = -1.0 * (*ny);
    }
    return;
  }
  if (circle_radius(bc) == 0.0) return;
  const double t = angle_of(bc, x);
  // codes 11, 13 and 14 point at the origin, the others away from it
  // (curved_boundaries.C:728-786, 842-862)
  if (bc == 11 || bc == 13 || bc == 14) {
    *nx = -std::cos(t);
    *ny = -std::sin(t);
  } else {
    *nx = std::cos(t);
    *ny = std::sin(t);
  }
}

double curved_arclength(int bc, const double* xL, const double* xR) {
  if (is_naca(bc)) {
    // curved_boundaries.C:1095-1140
    double xa, xb;
    if (xL[0] > xR[0]) { xa = xR[0]; xb = xL[0]; }
    else { xa = xL[0]; xb = xR[0]; }
    return naca_arclength(xa, xb);
  }
  const double R = circle_radius(bc);
  if (R == 0.0) return 0.0;
  double tL = angle_of(bc, xL), tR = angle_of(bc, xR);
  // codes 13-16 take the angle difference as it is (curved_boundaries.C:1063-1093)
  if (bc == 11 || bc == 12 || bc == 21 || bc == 22) same_side(&tL, &tR);
  if (bc == 11 || bc == 12) return std::fabs(tR - tL);
  return std::fabs(tR - tL) * R;
}

}  // namespace fs
