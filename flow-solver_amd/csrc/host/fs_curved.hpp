// Analytic curved walls (boundary codes above 10), curved_boundaries.C:28-1155.
//
//   11/12  unit circle at the origin (wall / farfield)
//   13/14  circle R=2               15/16  circle R=3
//   17/18  NACA 0012 upper surface  19/20  NACA 0012 lower surface
//   21/22  circle R=300
//
// Points are (x, y) pairs.  Every function keeps the reference's quirks: the
// circle angles are flipped to one side when a segment's end angles differ in
// sign (meshes put nodes on y = 0 so that no segment crosses the x axis), the
// NACA midpoint is the profile point above the chord midpoint (the
// reference returns before its arclength Newton iteration), and the normals
// point out of the fluid domain except for code 14, whose normal the
// reference points at the origin like code 13.
//
// The reference integrates the NACA arclength with GSL's QAGS at 1e-8
// absolute and relative tolerance.  GSL is not part of this build: the
// arclength here is the same integral, near the leading edge after the
// substitution x = t^2 (which removes the 1/sqrt(x) singularity), evaluated
// by adaptive Gauss-Legendre quadrature to about 1e-15 relative.  Results
// agree with the reference's to its quadrature tolerance, not bit for bit.
#pragma once
#include <string>

namespace fs {

// 11..22: the codes curved_boundaries.C:28-44 reserves
inline bool curved_bc(int bc) { return bc > 10; }
inline bool curved_bc_known(int bc) { return bc >= 11 && bc <= 22; }

// y of the profile at chord position x (naca_function, curved_boundaries.C:47-56)
double naca_y(double x, int bc);
// arclength of the profile between chord positions xa <= xb
double naca_arclength(double xa, double xb);

// curved_boundary_midpoint (curved_boundaries.C:91-330)
void curved_midpoint(int bc, const double* xL, const double* xR, double* xM);
// curved_boundary_gauss_points (curved_boundaries.C:345-690): the 3-point
// Gauss points of the arc from xL to xR, GP = (mid, +sqrt(3/5), -sqrt(3/5)).
// false (with *err) where the reference exits: Newton's iteration for a NACA
// point not converged in 50 steps.
bool curved_gauss_points(int bc, const double* xL, const double* xR, double* GP, std::string* err);
// curved_boundary_normal_vector (curved_boundaries.C:701-864)
void curved_normal(int bc, const double* x, double* nx, double* ny);
// curved_boundary_arclength (curved_boundaries.C:1009-1155)
double curved_arclength(int bc, const double* xL, const double* xR);

}  // namespace fs
