"""Analytic curved walls (boundary codes 11-22, curved_boundaries.C) on the
host: the primitives and the median-dual geometry built on them.

The reference's curved_boundaries.C does not compile here (it includes GSL,
which the image lacks), so there are no reference vectors for these codes.
They are pinned instead by what the mathematics fixes:

  * circles: midpoint, Gauss points, normals and arclength against the
    closed forms on the circle, including the reference's quirks (code 14's
    normal points at the origin; codes 13-16 take the angle difference
    without branch-cut handling);
  * the NACA arclength against scipy's QUADPACK QAGS (the algorithm GSL's
    gsl_integration_qags implements) at 1e-14 tolerance, and the NACA Gauss
    points at their arclength fractions;
  * whole meshes: the control volumes tile the curved domain, so the sums
    over nodes of area-weighted moments equal the domain's integrals of 1,
    x, y, x^2, y^2 and xy in closed form (annulus of circles; O-mesh of the
    analytic profile inside a polygonal farfield).  Dropping the subedge
    fix-up of Fix_Subedge_Data (geometry.C:852-981) moves these by 1e-7 to
    1e-5; with it they hold to 1e-13 of their scale;
  * a uniform state exerts no net force on a closed curved wall (the
    body-force Gauss normals integrate to zero around the profile).
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import fsutil as U

fsm = U.fsm

NACA_C = [0.2969, -0.1260, -0.3516, 0.2843, -0.1036]
NACA_E = [0.5, 1.0, 2.0, 3.0, 4.0]


def naca_y(x):
    return 0.6 * sum(c * x ** e for c, e in zip(NACA_C, NACA_E))


def naca_ds(x):
    yp = 0.08907 / math.sqrt(x) - 0.07560 - 0.42192 * x + 0.51174 * x * x - 0.24864 * x ** 3
    return math.sqrt(1.0 + yp * yp)


def qags(a, b):
    return quad(naca_ds, a, b, epsabs=1e-14, epsrel=1e-14, limit=500)[0]


def _params(order=3):
    return fsm.Params.make(order=order, foits=2, soits=2, limit=1, mach=0.5, alpha=1.0, CFL=20.0, ramp=10, subits=8,
                           method=1, iter=1, jufreq=1, reorder=1, gamma=1.4)


# ----------------------------------------------------------------- primitives


@pytest.mark.parametrize("a,b", [(0.0, 1.0), (0.0, 1e-4), (1e-6, 2e-6), (0.3, 0.31), (0.99, 1.0), (0.2, 0.7)])
def test_naca_arclength_matches_qags(a, b):
    ours = fsm.curved_arclength(17, (a, naca_y(a)), (b, naca_y(b)))
    rev = fsm.curved_arclength(19, (b, -naca_y(b)), (a, -naca_y(a)))
    ref = qags(a, b)
    assert abs(ours - ref) <= 1e-13 * ref
    assert ours == rev


def test_naca_arclength_random_intervals():
    """Many intervals, from a few 1e-12 chord to the whole chord: each
    matches QAGS, and the sum over a partition equals the whole."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        a = rng.uniform(0, 1) ** 3
        b = min(1.0, a + 10 ** rng.uniform(-12, 0))
        ours = fsm.curved_arclength(17, (a, naca_y(a)), (b, naca_y(b)))
        ref = qags(a, b)
        assert abs(ours - ref) <= 1e-13 * ref + 1e-17, (a, b, ours, ref)
    cuts = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, 63)), [1.0]])
    parts = sum(fsm.curved_arclength(17, (p0, naca_y(p0)), (p1, naca_y(p1))) for p0, p1 in zip(cuts[:-1], cuts[1:]))
    whole = fsm.curved_arclength(17, (0.0, 0.0), (1.0, naca_y(1.0)))
    assert abs(parts - whole) <= 1e-14 * whole


def test_naca_gauss_points_at_arclength_fractions():
    beta = [0.5, (1 + math.sqrt(0.6)) / 2, (1 - math.sqrt(0.6)) / 2]
    for bc, sgn in ((17, 1.0), (19, -1.0)):
        for xa, xb in ((0.0, 0.004), (0.1, 0.2), (0.95, 1.0)):
            # either end first: points ordered from the end of smaller x
            for xL, xR in ((xa, xb), (xb, xa)):
                gp = fsm.curved_gauss_points(bc, (xL, sgn * naca_y(xL)), (xR, sgn * naca_y(xR)))
                Lt = qags(xa, xb)
                for (x, y), bt in zip(gp, beta):
                    assert y == sgn * naca_y(x) or abs(y - sgn * naca_y(x)) <= 1e-16
                    assert abs(qags(xa, x) - bt * Lt) <= 1e-12
    # the reference's midpoint is the profile point above the chord midpoint
    xm = fsm.curved_midpoint(17, (0.1, naca_y(0.1)), (0.3, naca_y(0.3)))
    assert xm == (0.2, naca_y(0.2)) or abs(xm[1] - naca_y(0.2)) <= 1e-16


def test_naca_normals_unit_orthogonal_and_into_the_profile():
    for x in (1e-6, 0.01, 0.2, 0.2969, 0.5, 0.9, 1.0):
        yp = 0.08907 / math.sqrt(x) - 0.07560 - 0.42192 * x + 0.51174 * x * x - 0.24864 * x ** 3
        for bc, sgn in ((17, 1.0), (19, -1.0)):
            nx, ny = fsm.curved_normal(bc, (x, sgn * naca_y(x)))
            assert abs(nx * nx + ny * ny - 1.0) <= 1e-15
            assert abs(nx * 1.0 + ny * sgn * yp) <= 1e-14 * math.hypot(1.0, yp)  # normal to the tangent
            assert sgn * ny <= 0.0  # points into the airfoil: out of the fluid
    assert fsm.curved_normal(17, (0.0, 0.0)) == (1.0, 0.0)


@pytest.mark.parametrize("bc,R", [(11, 1.0), (12, 1.0), (13, 2.0), (14, 2.0), (15, 3.0), (16, 3.0), (21, 300.0),
                                  (22, 300.0)])
def test_circle_primitives_closed_form(bc, R):
    inward = bc in (11, 13, 14)  # code 14 points at the origin as code 13 does
    for tL, tR in ((0.3, 0.5), (-0.5, -0.2), (2.9, 3.1), (-3.1, -2.9), (0.0, 0.2), (-0.2, 0.0)):
        xL, xR = (R * math.cos(tL), R * math.sin(tL)), (R * math.cos(tR), R * math.sin(tR))
        tm, h = 0.5 * (tL + tR), 0.5 * (tR - tL)
        xm = fsm.curved_midpoint(bc, xL, xR)
        assert math.hypot(xm[0] - R * math.cos(tm), xm[1] - R * math.sin(tm)) <= 1e-14 * R
        gp = fsm.curved_gauss_points(bc, xL, xR)
        for (gx, gy), t in zip(gp, (0.0, math.sqrt(0.6), -math.sqrt(0.6))):
            th = tm + h * t
            assert math.hypot(gx - R * math.cos(th), gy - R * math.sin(th)) <= 1e-14 * R
        nx, ny = fsm.curved_normal(bc, xL)
        s = -1.0 if inward else 1.0
        assert math.hypot(nx - s * math.cos(tL), ny - s * math.sin(tL)) <= 1e-15
        assert abs(fsm.curved_arclength(bc, xL, xR) - R * abs(tR - tL)) <= 1e-14 * R
    # across the branch cut at the angle pi: codes 11/12 and 21/22 put both
    # ends on one side; codes 13-16 take the raw angle difference
    tL, tR = 3.1, -3.1
    xL, xR = (R * math.cos(tL), R * math.sin(tL)), (R * math.cos(tR), R * math.sin(tR))
    L = fsm.curved_arclength(bc, xL, xR)
    if bc in (13, 14, 15, 16):
        assert abs(L - R * 6.2) <= 1e-12 * R
    else:
        assert abs(L - R * 0.0) <= 1e-12 * R  # both ends flipped to -3.1


def test_unknown_curved_codes_rejected():
    with pytest.raises(fsm.FlowSolverError):
        fsm.curved_normal(23, (1.0, 0.0))
    with pytest.raises(fsm.FlowSolverError):
        fsm.Mesh.annulus(32, 5, 2.0, 3.0, bc_in=13, bc_out=16)
    with pytest.raises(fsm.FlowSolverError):
        fsm.Mesh.airfoil_split(32, 5, upper_bc=19, lower_bc=17)


# ------------------------------------------------------------------- meshes


def _domain_sums(m):
    """Sums over nodes of A, A x, A y, A x^2, A y^2, A xy rebuilt from the
    moments about each node (Moments / A are the CV means of (x-xi), (y-yi),
    (x-xi)^2, (y-yi)^2, (x-xi)(y-yi))."""
    A = m.array("cv_area")
    nn = len(A)
    M = m.array("moments").reshape(-1, 9)
    X = m.array("x").reshape(-1, 2)[:nn]
    x, y = X[:, 0], X[:, 1]
    Ix, Iy, Ixx, Iyy, Ixy = (M[:, k] for k in range(5))
    return dict(one=A.sum(), x=(A * (Ix + x)).sum(), y=(A * (Iy + y)).sum(),
                xx=(A * (Ixx + 2 * x * Ix + x * x)).sum(), yy=(A * (Iyy + 2 * y * Iy + y * y)).sum(),
                xy=(A * (Ixy + x * Iy + y * Ix + x * y)).sum())


def _polygon_integrals(P, Q):
    """Integrals of 1, x, y, x^2, y^2, xy over the polygon with directed edges P->Q (counter-clockwise)."""
    c = P[:, 0] * Q[:, 1] - Q[:, 0] * P[:, 1]
    x0, y0, x1, y1 = P[:, 0], P[:, 1], Q[:, 0], Q[:, 1]
    return dict(one=c.sum() / 2, x=((x0 + x1) * c).sum() / 6, y=((y0 + y1) * c).sum() / 6,
                xx=((x0 * x0 + x0 * x1 + x1 * x1) * c).sum() / 12, yy=((y0 * y0 + y0 * y1 + y1 * y1) * c).sum() / 12,
                xy=((x0 * y1 + 2 * x0 * y0 + 2 * x1 * y1 + x1 * y0) * c).sum() / 24)


@pytest.mark.parametrize("ntheta", [64, 128])
def test_annulus_of_circles_domain_integrals(ntheta):
    """Unit-circle wall (code 11) inside the R=300 farfield circle (code 22)."""
    m = fsm.Mesh.annulus(ntheta, 13, r_in=1.0, r_out=300.0, bc_in=11, bc_out=22, jitter=0.1, seed=7)
    m.prepare(_params())
    d = _domain_sums(m)
    R4 = math.pi * (300.0 ** 4 - 1.0) / 4
    exact = dict(one=math.pi * (300.0 ** 2 - 1.0), x=0.0, y=0.0, xx=R4, yy=R4, xy=0.0)
    scale = dict(one=300.0 ** 2, x=300.0 ** 3, y=300.0 ** 3, xx=300.0 ** 4, yy=300.0 ** 4, xy=300.0 ** 4)
    for k in exact:
        print(k, d[k], exact[k], (d[k] - exact[k]) / scale[k])
        assert abs(d[k] - exact[k]) <= 1e-12 * scale[k], k


@pytest.mark.parametrize("nwall", [128, 256])
def test_naca_omesh_domain_integrals(nwall):
    """Analytic upper/lower surfaces (codes 17/19) inside a polygonal farfield:
    domain = polygon minus the profile, whose integrals are closed forms in
    the profile's coefficients (y^2 by quadrature of y^3)."""
    m = fsm.Mesh.airfoil_split(nwall, 17, upper_bc=17, lower_bc=19, h0=1e-3, rmax=5.0, jitter=0.1, seed=3)
    m.prepare(_params())
    d = _domain_sums(m)
    X = m.array("x").reshape(-1, 2)
    off, seg = m.array("bseg_off"), m.array("bseg").reshape(-1, 2)
    far = seg[off[2]:off[3]]
    poly = _polygon_integrals(X[far[:, 0]], X[far[:, 1]])

    def prof(k):  # integral over the profile of x^k (y from -y(x) to y(x))
        return 2 * 0.6 * sum(c / (e + k + 1) for c, e in zip(NACA_C, NACA_E))
    yy_prof = quad(lambda x: 2.0 / 3.0 * naca_y(x) ** 3, 0, 1, epsabs=1e-15, epsrel=1e-14, limit=200)[0]
    exact = dict(one=poly["one"] - prof(0), x=poly["x"] - prof(1), y=poly["y"], xx=poly["xx"] - prof(2),
                 yy=poly["yy"] - yy_prof, xy=poly["xy"])
    for k in exact:
        sc = 5.0 ** (2 + {"one": 0, "x": 1, "y": 1}.get(k, 2))
        print(k, d[k], exact[k], (d[k] - exact[k]) / sc)
        assert abs(d[k] - exact[k]) <= 1e-13 * sc, k


def test_curved_boundary_geometry_structure():
    """Ghosts at the curved midpoint, bedge lengths along the curve, the
    boundary subedge ending at the curved midpoint and normal to the line to
    the centroid, Gauss points on the curve with the curve's normals."""
    m = fsm.Mesh.airfoil_split(64, 9, upper_bc=17, lower_bc=19, h0=2e-2, rmax=8.0, jitter=0.1, seed=3)
    m.prepare(_params())
    info = m.info()
    nn = info["nn"]
    X = m.array("x").reshape(-1, 2)
    be = m.array("bedge").reshape(-1, 5)
    bn = m.array("bedge_n").reshape(-1, 3)
    bxg = m.array("bedge_xg").reshape(-1, 3, 2)
    bgn = m.array("bedge_gn").reshape(-1, 3, 2)
    bbc = m.array("bbc")
    off, seg = m.array("bseg_off"), m.array("bseg").reshape(-1, 2)
    edges = m.array("edges").reshape(-1, 2)
    sub_edge, sub_elem = m.array("sub_edge"), m.array("sub_elem")
    sub_n, sub_xm = m.array("sub_n").reshape(-1, 3), m.array("sub_xm").reshape(-1, 2)
    cent = m.array("el_cent").reshape(-1, 2)
    ncurved = 0
    for i in range(info["nbe"]):
        node, ghost, g, b, s = be[i]
        bc = bbc[b]
        if bc not in (17, 19):
            np.testing.assert_array_equal(bgn[i], np.tile(bn[i, :2], (3, 1)))
            continue
        ncurved += 1
        n0, n1 = seg[off[b] + s]
        xm = fsm.curved_midpoint(bc, X[n0], X[n1])
        assert tuple(X[nn + ghost]) == xm
        assert bn[i, 0:2].tolist() == list(fsm.curved_normal(bc, X[node]))
        assert bn[i, 2] == fsm.curved_arclength(bc, X[node], X[nn + ghost])
        sgn = 1.0 if bc == 17 else -1.0
        for k in range(3):
            gx, gy = bxg[i, k]
            assert abs(gy - sgn * naca_y(gx)) <= 1e-16
            assert bgn[i, k].tolist() == list(fsm.curved_normal(bc, (gx, gy)))
        lo, hi = min(n0, n1), max(n0, n1)
        e = int(np.nonzero((edges[:, 0] == lo) & (edges[:, 1] == hi))[0][0])
        j = int(np.nonzero(sub_edge == e)[0][0])
        assert sub_elem[j] == g
        assert tuple(sub_xm[j]) == xm
        t = cent[g] - np.array(xm)
        assert abs(sub_n[j, 2] - np.hypot(*t)) <= 1e-15 * sub_n[j, 2]
        assert abs(sub_n[j, 0] * t[0] + sub_n[j, 1] * t[1]) <= 1e-15 * sub_n[j, 2]
    assert ncurved == 2 * 64


def test_uniform_state_no_force_on_closed_curved_wall():
    """Uniform pressure on a closed wall: zero net force.  The curved branch
    of Compute_Body_Forces (grid_io.C:1612-1660) integrates the curve's
    normals at its Gauss points with weights 8/9, 5/9, 5/9; the polyline
    branch its segment normals.  A wrong weight order leaves an O(h)
    residue; the curved sum closes to quadrature accuracy."""
    out = {}
    for bcs in ((17, 19), (1, 1)):
        m = fsm.Mesh.airfoil_split(256, 9, upper_bc=bcs[0], lower_bc=bcs[1], h0=2e-2, rmax=8.0, jitter=0.1, seed=3)
        p = _params(order=1)
        m.prepare(p)
        o = U.Oracle(m, p)
        o.init_freestream()
        f = o.body_forces(1)
        out[bcs] = f
        print(bcs, f)
    assert max(abs(v) for v in out[(1, 1)]) <= 1e-15
    # measured 6.6e-13 with the weights 8/9, 5/9, 5/9; 4e-4 with 5/9, 8/9, 5/9
    assert max(abs(v) for v in out[(17, 19)]) <= 1e-11


def test_curved_codes_through_the_grid_files(tmp_path):
    """A mesh with curved codes written in the reference's grid and bc.dat
    formats and read back prepares to the same curved geometry."""
    m = fsm.Mesh.airfoil_split(64, 9, upper_bc=17, lower_bc=19, h0=2e-2, rmax=8.0, jitter=0.1, seed=3)
    g, bc = str(tmp_path / "naca.mesh"), str(tmp_path / "bc.dat")
    m.write(g, bc)
    assert open(bc).read().split() == ["17", "19", "0"]
    m2 = fsm.Mesh.read(g, bc)
    p = _params()
    m.prepare(p)
    m2.prepare(p)
    for name in ("x", "bedge_n", "bedge_xg", "bedge_gn", "sub_n", "sub_xm", "cv_area", "moments"):
        np.testing.assert_array_equal(m.array(name), m2.array(name), err_msg=name)


@pytest.mark.parametrize("ntheta", [64, 128])
def test_quad_annulus_of_circles_domain_integrals(ntheta):
    """Quadrilateral cells on the curved walls: the boundary subedge of a
    curved segment then starts at a quadrilateral's centroid.  The same
    closed-form domain integrals as on the triangle annulus."""
    m = U.annulus_quads(ntheta, 13)
    m.prepare(_params())
    assert m.info()["ntri"] == 0 and m.info()["nquad"] == 12 * ntheta
    d = _domain_sums(m)
    R4 = math.pi * (300.0 ** 4 - 1.0) / 4
    exact = dict(one=math.pi * (300.0 ** 2 - 1.0), x=0.0, y=0.0, xx=R4, yy=R4, xy=0.0)
    scale = dict(one=300.0 ** 2, x=300.0 ** 3, y=300.0 ** 3, xx=300.0 ** 4, yy=300.0 ** 4, xy=300.0 ** 4)
    for k in exact:
        print(k, d[k], exact[k], (d[k] - exact[k]) / scale[k])
        assert abs(d[k] - exact[k]) <= 1e-12 * scale[k], k
