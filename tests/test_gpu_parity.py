"""GPU parity: the HIP path through the C-ABI against the oracle (the C
restatement pinned bit for bit to the reference build) on the same seeded
inputs.

Tolerances (FP64 path):
  * face fluxes, residuals, gradients, limiter, ev_con, dt: per-node residual
    error <= 1e-12 normalised by the sum of |face flux| at the node; in
    practice these agree bit for bit because the device keeps the reference's
    operation order and gathers in ascending face order;
  * Jacobian blocks: off-diagonal blocks bit-exact (same FD arithmetic);
    diagonal blocks 1e-12 relative (different summation order);
  * one pseudo-time step from an identical state: dQ within 1e-10 relative of
    the oracle's multicolor sweep (the diagonal and each row product are summed in
    a different fixed order, 1e-16 apart).  Trajectories are not compared
    step by step: the reference's forward-difference Jacobian (eps=1e-8)
    amplifies any last-bit state difference by 1e8;
  * converged steady state: density L2 relative difference <= 1e-10.
"""
import numpy as np
import pytest

import fsutil as U

fsm = U.fsm
pytestmark = pytest.mark.gpu


def _params(**kw):
    base = dict(nsteps=50, iter=1, jufreq=1, CFL=20.0, ramp=10, gamma=1.4, alpha=1.0, mach=0.5, subits=10, method=1,
                order=1, foits=2, soits=0, limit=1, reorder=1)
    base.update(kw)
    return fsm.Params.make(**base)


def _mesh(kind, nx, ny, seed):
    """bump: channel of nx x ny nodes; airfoil: O-mesh of nx wall x ny radial
    nodes around the 12%-thick profile (polyline wall); naca: the same with the
    analytic NACA 0012 upper/lower surfaces (codes 17/19); cylinder: unit-circle
    wall (code 11) in the R=300 farfield circle (code 22); mixed: the bump
    channel with a third of its cells quadrilaterals; cylquad: the cylinder on
    quadrilaterals only."""
    if kind == "bump":
        return fsm.Mesh.bump(nx, ny, 0.2, seed)
    if kind == "airfoil":
        return fsm.Mesh.airfoil(nx, ny, 2e-2, 8.0, 0.1, seed)
    if kind == "naca":
        return fsm.Mesh.airfoil_split(nx, ny, upper_bc=17, lower_bc=19, h0=2e-2, rmax=8.0, jitter=0.1, seed=seed)
    if kind == "cylinder":
        return fsm.Mesh.annulus(nx, ny, r_in=1.0, r_out=300.0, bc_in=11, bc_out=22, jitter=0.1, seed=seed)
    if kind == "mixed":
        return U.bump_mixed(nx, ny, seed)
    if kind == "cylquad":
        return U.annulus_quads(nx, ny)
    raise ValueError(kind)


def _pair(nx, ny, seed, airfoil=False, kind=None, **kw):
    """Device solver and oracle on one seeded mesh (see _mesh; airfoil=True is kind "airfoil")."""
    p = _params(**kw)
    m = _mesh(kind or ("airfoil" if airfoil else "bump"), nx, ny, seed)
    m.prepare(p)
    s = fsm.Solver(m, p, fsm.FS_PREC_F64)
    o = U.Oracle(m, p)
    return p, m, s, o


def _set_state(s, o, Q):
    s.set("Q", Q)
    s.set("nQ", Q)
    o["Q"][:] = Q
    o["nQ"][:] = Q


@pytest.mark.parametrize("ts,order,soits,limit", [(1, 1, 0, 1), (5, 1, 0, 1), (4, 3, 4, 1), (5, 1, 0, 2)])
def test_residual_and_jacobian_mirrors(ts, order, soits, limit):
    p, m, s, o = _pair(33, 17, 11, order=order, soits=soits, limit=limit)
    info = m.info()
    Q = U.perturbed_state(info["nn"], info["ng"], p, seed=5)
    _set_state(s, o, Q)
    o.s.citer = ts
    # reference call sequence of solve.C:352-767
    s.build_matrix(ts)
    assert U.oracle_call(o, "fso_build_matrix") == 0
    phase_grad = (order == 1 and ts > p.foits) or (order == 3 and p.foits < ts <= p.foits + p.soits)
    if phase_grad:
        s.gradient()
        U.oracle_call(o, "fso_gradient")
        np.testing.assert_array_equal(s.get("grad"), o["grad"])
    if order == 1 and ts > p.foits:
        s.limiter(ts)
        U.oracle_call(o, "fso_limiter_venkat" if limit == 2 else "fso_limiter_barth")
        np.testing.assert_array_equal(s.get("phi"), o["phi"][0::8])
    s.interior_residual(1, ts)
    U.oracle_call(o, "fso_interior_residual", 1, ts)
    s.boundary_residual(1)
    assert U.oracle_call(o, "fso_boundary_residual", 1) == 0
    np.testing.assert_array_equal(s.get("ev_con"), o["ev_con"])
    np.testing.assert_array_equal(s.get("dt"), o["dt"])
    R, Rr = s.get("R"), o["R"]
    err = np.max(np.abs(R - Rr)) / np.max(np.abs(Rr))
    print("ts=%d order=%d residual rel err %.3e bitexact=%s" % (ts, order, err, np.array_equal(R, Rr)))
    assert err <= 1e-12
    np.testing.assert_array_equal(s.get("nQ"), o["nQ"])  # ghost states with memory
    L, Lr = s.get("LHS").reshape(-1, 16), o["LHS"].reshape(-1, 16)
    iau = m.array("iau")
    off = np.ones(L.shape[0], bool)
    off[iau] = False
    np.testing.assert_array_equal(L[off], Lr[off])
    derr = np.max(np.abs(L[iau] - Lr[iau])) / np.max(np.abs(Lr[iau]))
    print("diag block rel err %.3e" % derr)
    assert derr <= 1e-12


def test_roe_batch_bitexact():
    import ctypes as C
    rng = np.random.default_rng(3)
    n = 4096
    gamma = 1.4
    th = rng.uniform(0, 2 * np.pi, n)
    nrm = np.stack([np.cos(th), np.sin(th), rng.uniform(0.01, 1.0, n)], 1).ravel()
    def states():
        rho = rng.uniform(0.5, 2.0, n)
        u = rng.uniform(-1.5, 1.5, n)
        v = rng.uniform(-1.5, 1.5, n)
        pr = rng.uniform(0.3, 2.0, n)
        return np.stack([rho, rho * u, rho * v, pr / (gamma - 1) + 0.5 * rho * (u * u + v * v)], 1).ravel()
    qL, qR = states(), states()
    # device buffers through hipMalloc via torch-free ctypes: reuse a solver-less path
    hip = C.CDLL("libamdhip64.so")
    ptrs = []
    for arr in (nrm, qL, qR, np.zeros(4 * n)):
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), C.c_size_t(arr.nbytes)) == 0
        assert hip.hipMemcpy(d, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
        ptrs.append(d)
    rc = fsm.lib().fs_roe_flux_batch(n, gamma, ptrs[0], ptrs[1], ptrs[2], ptrs[3], None)
    assert rc == 0
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(4 * n)
    assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptrs[3], C.c_size_t(out.nbytes), 2) == 0
    for d in ptrs:
        hip.hipFree(d)
    ref = np.concatenate([U.roe_oracle(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], gamma, qL[4 * i:4 * i + 4],
                                       qR[4 * i:4 * i + 4]) for i in range(n)])
    assert np.array_equal(out, ref), np.max(np.abs(out - ref))


def test_roe_wave_form_matches_reference_form():
    """The characteristic-wave Roe flux of the mixed path's FP64 residual
    against the oracle's restatement of flux.C:299-541: per face, within 1e-13
    of the sum of the magnitudes of the two physical fluxes (the terms the
    flux is formed from)."""
    import ctypes as C
    rng = np.random.default_rng(4)
    n = 8192
    gamma = 1.4
    th = rng.uniform(0, 2 * np.pi, n)
    nrm = np.stack([np.cos(th), np.sin(th), rng.uniform(0.01, 1.0, n)], 1).ravel()

    def states():
        rho = rng.uniform(0.3, 3.0, n)
        u = rng.uniform(-2.0, 2.0, n)
        v = rng.uniform(-2.0, 2.0, n)
        pr = rng.uniform(0.2, 3.0, n)
        return np.stack([rho, rho * u, rho * v, pr / (gamma - 1) + 0.5 * rho * (u * u + v * v)], 1).ravel()

    qL, qR = states(), states()
    hip = C.CDLL("libamdhip64.so")
    ptrs = []
    for arr in (nrm, qL, qR, np.zeros(4 * n)):
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), C.c_size_t(arr.nbytes)) == 0
        assert hip.hipMemcpy(d, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
        ptrs.append(d)
    lib = fsm.lib()
    lib.fs_roe_flux_batch_form.argtypes = [C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 5
    assert lib.fs_roe_flux_batch_form(n, 1, gamma, ptrs[0], ptrs[1], ptrs[2], ptrs[3], None) == 0
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(4 * n)
    assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptrs[3], C.c_size_t(out.nbytes), 2) == 0
    for d in ptrs:
        hip.hipFree(d)
    worst = 0.0
    for i in range(n):
        nx, ny, ln = nrm[3 * i:3 * i + 3]
        ref = U.roe_oracle(nx, ny, ln, gamma, qL[4 * i:4 * i + 4], qR[4 * i:4 * i + 4])
        scale = U.roe_oracle(nx, ny, ln, gamma, qL[4 * i:4 * i + 4], qL[4 * i:4 * i + 4])
        scale = np.abs(scale) + np.abs(U.roe_oracle(nx, ny, ln, gamma, qR[4 * i:4 * i + 4], qR[4 * i:4 * i + 4]))
        worst = max(worst, float(np.max(np.abs(out[4 * i:4 * i + 4] - ref) / (scale + 1e-300))))
    print("wave-form Roe flux: max error / (|F(qL)| + |F(qR)|) = %.2e" % worst)
    assert worst <= 1e-13


@pytest.mark.parametrize("order,soits,limit,airfoil,newton", [
    (1, 0, 1, False, (1, 1)), (3, 2, 1, False, (1, 1)), (1, 0, 2, False, (1, 1)), (3, 2, 2, False, (1, 1)),
    (1, 0, 1, True, (1, 1)), (3, 2, 1, True, (1, 1)),
    # analytic curved walls (curved_boundaries.C)
    (1, 0, 1, "naca", (1, 1)), (3, 2, 1, "naca", (1, 1)), (1, 0, 1, "cylinder", (1, 1)),
    (3, 2, 1, "cylinder", (1, 1)),
    # triangles and quadrilaterals; quadrilaterals on curved walls
    (1, 0, 1, "mixed", (1, 1)), (3, 2, 1, "mixed", (1, 1)), (3, 2, 1, "cylquad", (1, 1)),
    # 3 Newton iterations per step with the Jacobian updated every 2nd
    # (solve.C:328-338): iterations without an update keep only V/dt
    (1, 0, 1, False, (3, 2)), (3, 2, 1, False, (2, 2))])
def test_step_history(order, soits, limit, airfoil, newton):
    """fs_solve against the oracle's solve with the same multicolor sweep order
    (bump channel, the O-mesh with its closed polyline wall, the O-mesh with
    the analytic NACA surfaces, the cylinder in a circular farfield; several
    Newton iterations per step with Jacobian reuse).

    dQ bar: 1e-10 with one Newton iteration per step.  With several, the
    sweeps' last-bit differences (1e-17 relative, from the summation order of
    the diagonal and row products) reach the next iteration's state inside
    the step, and the forward-difference Jacobian (eps = 1e-8) and the
    limiter scale them up (measured: dQ 1.8e-10 at a first-order step, 1.3e-8
    at the first limited step): 1e-6 then.  Each iteration is pinned from an
    identical start by test_newton_iterations_match_oracle."""
    it, ju = newton
    dq_bar = 1e-10 if it == 1 else 1e-6
    if airfoil == "mixed":
        p, m, s, o = _pair(33, 17, 2, kind="mixed", order=order, soits=soits, limit=limit, subits=10)
    elif airfoil in ("naca", "cylinder", "cylquad"):
        p, m, s, o = _pair(40, 11, 4, kind=airfoil, order=order, soits=soits, limit=limit, subits=10, alpha=2.0,
                           mach=0.5 if airfoil == "naca" else 0.3)
    elif airfoil:
        p, m, s, o = _pair(40, 11, 4, airfoil=True, order=order, soits=soits, limit=limit, subits=10, alpha=2.0)
    else:
        p, m, s, o = _pair(33, 17, 2, order=order, soits=soits, limit=limit, subits=10, iter=it, jufreq=ju)
    s.init_freestream()
    o.init_freestream()
    nsteps = p.foits + soits + 4
    for ts in range(1, nsteps + 1):
        # re-synchronise: the FD Jacobian (eps=1e-8) turns last-bit state
        # differences into 1e-8 Jacobian differences, so each step is checked
        # from an identical start
        o["Q"][:] = s.get("Q")
        dq, r2 = s.solve(ts)
        dqr = o.solve(ts, multicolor=1)
        R, Rr = s.get("R"), o["R"]
        rerr = np.max(np.abs(R - Rr)) / np.max(np.abs(Rr))
        dQ, dQr = s.get("dQ"), o["dQ"]
        derr = np.max(np.abs(dQ - dQr)) / np.max(np.abs(dQr))
        print("step %d |dQ| %.15e vs %.15e  R err %.2e (bitexact %s)  dQ err %.2e" %
              (ts, dq, dqr, rerr, np.array_equal(R, Rr), derr))
        assert rerr <= 1e-12
        assert derr <= dq_bar
        assert abs(dq - dqr) <= dq_bar * abs(dqr)
    if order == 3:
        hr = o["hess"]
        herr = np.max(np.abs(s.get("hess") - hr)) / np.max(np.abs(hr))
        print("k=2 derivative rel err vs oracle (independent pseudo-inverse) %.2e" % herr)
        assert herr <= 1e-8


def test_converged_state_k1_unlimited():
    """Steady state to a 1e-10 residual drop: density L2 within 1e-10."""
    p, m, s, o = _pair(49, 17, 1, order=1, limit=0, CFL=100.0, ramp=30, foits=10, subits=20, alpha=0.0)
    s.init_freestream()
    o.init_freestream()
    steps, drop = s.solve_steady(600, 1e-10)
    r0 = None
    for ts in range(1, 601):
        o.solve(ts, multicolor=0)
        r = np.sqrt(o.s.last_rms)
        if ts == p.foits + 1:
            r0 = r
        if r0 and r / r0 <= 1e-10:
            break
    nn = m.info()["nn"]
    rho = s.get("Q")[0:4 * nn:4]
    rho_ref = o["Q"][0:4 * nn:4]
    rel = np.linalg.norm(rho - rho_ref) / np.linalg.norm(rho_ref)
    print("gpu steps %d drop %.2e; oracle steps %d; density L2 rel diff %.3e" % (steps, drop, ts, rel))
    assert drop <= 1e-10
    assert rel <= 1e-10


def test_mixed_fp32_jacobian_blocks():
    """FP32 analytic blocks of the mixed path against the oracle's FP64
    restatement of jacobian.C:301-658 at the same (limited, reconstructed) states."""
    p = _params(order=1, foits=2, subits=0)
    m = fsm.Mesh.bump(33, 17, 0.2, 11)
    m.prepare(p)
    s = fsm.Solver(m, p, fsm.FS_PREC_MIXED)
    o = U.Oracle(m, p)
    info = m.info()
    Q = U.perturbed_state(info["nn"], info["ng"], p, seed=8)
    _set_state(s, o, Q)
    ts = 5
    s.solve(ts)
    U.oracle_lib().fso_set_analytic_jacobian(1)
    try:
        o.solve(ts, multicolor=1)
    finally:
        U.oracle_lib().fso_set_analytic_jacobian(0)
    L32, Lr = s.get("LHS32").reshape(-1, 16), o["LHS"].reshape(-1, 16)
    err = np.max(np.abs(L32 - Lr)) / np.max(np.abs(Lr))
    print("FP32 Jacobian blocks vs FP64 analytic: max rel err %.2e" % err)
    assert err <= 2e-5
    R, Rr = s.get("R"), o["R"]
    assert np.max(np.abs(R - Rr)) <= 1e-12 * np.max(np.abs(Rr))


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED, fsm.FS_PREC_MIXED16])
@pytest.mark.parametrize("order,limit,soits", [(1, 0, 0), (3, 1, 10)])
def test_mixed_converged_state_matches_fp64(order, limit, soits, precision):
    """Mixed precision (FP32 sweep operands, or FP16 ones) reaches the same
    steady state as the FP64 path: the FP64 outer residual fixes the
    converged state; error against the oracle stated."""
    p, m, s, o = _pair(49, 17, 1, order=order, limit=limit, soits=soits, CFL=100.0, ramp=30, foits=10, subits=20,
                       alpha=0.0)
    sm = fsm.Solver(m, p, precision)
    sm.init_freestream()
    steps_m, drop_m = sm.solve_steady(800, 1e-10)
    o.init_freestream()
    ts_ref = p.foits + 1 if order == 1 else p.foits + soits + 1
    r0 = None
    for ts in range(1, 801):
        o.solve(ts, multicolor=0)
        r = np.sqrt(o.s.last_rms)
        if ts == ts_ref:
            r0 = r
        if r0 and r / r0 <= 1e-10:
            break
    nn = m.info()["nn"]
    rho, rho_ref = sm.get("Q")[0:4 * nn:4], o["Q"][0:4 * nn:4]
    rel = np.linalg.norm(rho - rho_ref) / np.linalg.norm(rho_ref)
    print("mixed steps %d drop %.2e vs oracle %d steps; density L2 rel err vs FP64 oracle %.3e" %
          (steps_m, drop_m, ts, rel))
    assert drop_m <= 1e-10
    assert rel <= 1e-10


def _oracle_from_device(s, o, info):
    """The oracle's state and reconstruction fields set to the device's."""
    nn = info["nn"]
    o["nQ"][:] = s.get("nQ")
    o["Q"][:] = s.get("Q")
    o["grad"][:] = s.get("grad")
    o["phi"][:] = np.repeat(s.get("phi"), 8)
    o["hess"][:] = s.get("hess")
    o["pv"][:] = s.get("pv")
    assert o["pv"].size == 4 * nn


@pytest.mark.parametrize("order,soits,limit,airfoil", [(1, 0, 1, False), (3, 2, 1, False), (3, 2, 2, False),
                                                       (1, 0, 0, False), (3, 2, 1, True), (3, 2, 1, "naca"),
                                                       (3, 2, 1, "cylinder"), (3, 2, 1, "cylquad")])
def test_body_forces_match_oracle(order, soits, limit, airfoil):
    """Compute_Body_Forces on the device (per-bedge Gauss-point terms summed
    in bedge order) equals the oracle's bit for bit from the same fields, in
    every phase: nodal, limited linear and the k=2 point-value quadratic; on
    curved walls with the curve's Gauss points and normals."""
    if airfoil in ("naca", "cylinder", "cylquad"):
        p, m, s, o = _pair(40, 11, 21, kind=airfoil, order=order, soits=soits, limit=limit, alpha=2.0, subits=8,
                           mach=0.5 if airfoil == "naca" else 0.3)
    elif airfoil:
        p, m, s, o = _pair(40, 11, 21, airfoil=True, order=order, soits=soits, limit=limit, alpha=2.0, subits=8)
    else:
        p, m, s, o = _pair(33, 17, 21, order=order, soits=soits, limit=limit, alpha=2.0, subits=8)
    info = m.info()
    s.init_freestream()
    nsteps = p.foits + soits + 3
    seen = set()
    for ts in range(1, nsteps + 1):
        s.solve(ts)
        _oracle_from_device(s, o, info)
        fd, fo = s.body_forces(ts), o.body_forces(ts)
        seen.add(0 if ts <= p.foits else (1 if (order == 1 or ts <= p.foits + soits) else 3))
        print("ts %d forces device %s oracle %s" % (ts, fd, fo), flush=True)
        np.testing.assert_array_equal(fd, fo)
    assert seen == ({0, 1} if order == 1 else {0, 1, 3})
    # surface Cp of every boundary, and the reference's file format
    for b in range(info["nb"]):
        xd, cd = s.surface_cp(b)
        xo, co = o.surface_cp(b)
        np.testing.assert_array_equal(xd, xo)
        np.testing.assert_array_equal(cd, co)


def test_surface_cp_file(tmp_path):
    p, m, s, o = _pair(17, 9, 22, order=1, limit=0)
    s.init_freestream()
    for ts in range(1, 4):
        s.solve(ts)
    s.write_surface_cp(str(tmp_path))
    bbc = m.array("bbc")
    for b in range(m.info()["nb"]):
        path = tmp_path / ("surface_%d_pressure.dat" % (b + 1))
        if bbc[b] == 0 or bbc[b] % 2 == 0:
            assert not path.exists()
            continue
        rows = [tuple(map(float, ln.split())) for ln in path.read_text().splitlines()]
        x, cp = s.surface_cp(b)
        assert len(rows) == len(x)
        xs = [r[0] for r in rows]
        assert xs == sorted(xs)
        want = sorted(zip(np.float32(x).tolist(), np.float32(cp).tolist()))
        np.testing.assert_allclose(np.array(rows), np.array(want), atol=1e-6)


@pytest.mark.parametrize("kind", ["airfoil", "naca"])
def test_transonic_airfoil_fp64_and_mixed_vs_oracle(kind):
    """The precision configuration in miniature: an O-mesh of 64 x 17 nodes
    around the 12%-thick profile (first cell 1e-3 chord, 30 chords, +-0.1
    jitter), M=0.8 at 1.25 degrees, k=2 with the Barth limiter; the wall a
    polyline (airfoil) or the analytic NACA 0012 surfaces (naca, codes 17/19,
    128 wall nodes).

    Fixed point: FP64 and mixed solves converged to a 1e-12 residual drop
    against the oracle's (same drop): density L2 within 1e-10, lift and drag
    (Compute_Body_Forces) within 1e-9 relative.  At the serving stopping test
    (1e-10 drop) a state is still a few 1e-10 from the fixed point on this
    case (the FP64 path 7e-10, measured), so that comparison is printed, and
    bounded only by 1e-8."""
    p = _params(order=3, foits=10, soits=10, limit=1, mach=0.8, alpha=1.25, CFL=100.0, ramp=30, subits=20)
    # the curved wall at 128 wall nodes: at 64, the arc's bulge at the leading
    # edge (1.8e-3) exceeds the first cell (1e-3) and the multicolor and
    # sequential sweeps both stop at a non-physical state with these controls
    m = (fsm.Mesh.airfoil(64, 17, 1e-3, 30.0, 0.1, 3) if kind == "airfoil" else
         fsm.Mesh.airfoil_split(128, 17, upper_bc=17, lower_bc=19, h0=1e-3, rmax=30.0, jitter=0.1, seed=3))
    m.prepare(p)
    nn = m.info()["nn"]
    o = U.Oracle(m, p)
    o.init_freestream()
    ts_ref, r0 = p.foits + p.soits + 1, None
    for ts in range(1, 2001):
        o.solve(ts, multicolor=0)
        r = np.sqrt(o.s.last_rms)
        if ts == ts_ref:
            r0 = r
        if r0 and r / r0 <= 1e-12:
            break
    f_ref = o.body_forces(ts)
    rho_ref = o["Q"][0:4 * nn:4]
    print("oracle: %d steps to a 1e-12 drop, forces %s" % (ts, f_ref), flush=True)
    for prec, name in ((fsm.FS_PREC_F64, "fp64"), (fsm.FS_PREC_MIXED, "mixed")):
        for tol in (1e-10, 1e-12):
            s = fsm.Solver(m, p, prec)
            s.init_freestream()
            steps, drop = s.solve_steady(2000, tol)
            f = s.body_forces(steps)
            rho = s.get("Q")[0:4 * nn:4]
            rel = np.linalg.norm(rho - rho_ref) / np.linalg.norm(rho_ref)
            ferr = np.abs(f - f_ref) / np.abs(f_ref)
            print("%s to %.0e: %d steps drop %.2e; density L2 rel err vs oracle %.3e; lift/drag rel err %.3e %.3e"
                  % (name, tol, steps, drop, rel, ferr[0], ferr[1]), flush=True)
            assert drop <= tol
            if tol == 1e-12:
                assert rel <= 1e-10
                assert np.all(ferr <= 1e-9)
            else:
                assert rel <= 1e-8
            s.close()


@pytest.mark.parametrize("limit", [1, 2])
def test_k2_gauss_limiter_matches_oracle(limit):
    """The k=2 phase limiter (Barth, Venkatakrishnan) at the Gauss points on
    the device against the oracle from identical inputs: phi bit for bit."""
    p, m, s, o = _pair(33, 17, 23, order=3, soits=2, limit=limit)
    info = m.info()
    Q = U.perturbed_state(info["nn"], info["ng"], p, seed=9, amp=0.08)
    _set_state(s, o, Q)
    k2_ts = p.foits + p.soits + 1
    s.derivatives()
    o["hess"][:] = s.get("hess")
    o["pv"][:] = s.get("pv")
    s.limiter(k2_ts)
    o.s.citer = k2_ts
    U.oracle_call(o, "fso_limiter_venkat" if limit == 2 else "fso_limiter_barth")
    phi_d, phi_o = s.get("phi"), o["phi"][0::8]
    print("k=2 limiter %d: min phi %.4f, %d of %d nodes limited" % (limit, phi_d.min(), int((phi_d < 1).sum()),
                                                                   phi_d.size))
    assert (phi_d < 1).sum() > 0
    np.testing.assert_array_equal(phi_d, phi_o)


def test_hbm_probe_in_range():
    """The bandwidth probe the bench reports beside the roofline: a streaming
    read and a copy of 1 GiB land between 1 TB/s and the 8 TB/s peak."""
    rd, cp = fsm.hbm_probe(1 << 30, 3)
    print("probe: read %.0f GB/s, copy %.0f GB/s" % (rd, cp))
    assert 1000.0 < rd < 8000.0
    assert 1000.0 < cp < 8000.0


@pytest.mark.parametrize("order,soits,ts", [(1, 0, 1), (1, 0, 3), (3, 2, 3), (3, 2, 5)])
def test_newton_iterations_match_oracle(order, soits, ts):
    """Three Newton iterations of one pseudo-time step with the Jacobian
    updated on the 1st and 3rd (solve.C:328-338: an iteration without an
    update keeps only V/dt on the diagonal), each started from the same state
    on the device and in the oracle: residual, off-diagonal blocks and ghost
    states bit for bit, diagonal blocks to 1e-12, dQ to 1e-10."""
    p, m, s, o = _pair(33, 17, 2, order=order, soits=soits, limit=1, iter=3, jufreq=2)
    info = m.info()
    nn = info["nn"]
    Q = U.perturbed_state(nn, info["ng"], p, seed=31)
    _set_state(s, o, Q)
    o.s.citer = ts
    iau = m.array("iau")
    phase_grad = (order == 1 and ts > p.foits) or (order == 3 and p.foits < ts <= p.foits + p.soits)
    k2 = order == 3 and ts > p.foits + p.soits
    lim = (order == 1 and ts > p.foits) or k2
    for it in (1, 2, 3):
        jac = 1 if (it - 1) % p.jufreq == 0 else 0
        o["phi"][:] = 1.0
        s.set("phi", np.ones(nn))
        s.build_matrix(ts)
        assert U.oracle_call(o, "fso_build_matrix") == 0
        if phase_grad:
            s.gradient()
            U.oracle_call(o, "fso_gradient")
        if k2:
            s.derivatives()
            U.oracle_call(o, "fso_derivatives")
            o["hess"][:] = s.get("hess")
            o["pv"][:] = s.get("pv")
        if lim:
            s.limiter(ts)
            U.oracle_call(o, "fso_limiter_barth")
            np.testing.assert_array_equal(s.get("phi"), o["phi"][0::8])
        s.interior_residual(jac, ts)
        U.oracle_call(o, "fso_interior_residual", jac, ts)
        s.boundary_residual(jac)
        assert U.oracle_call(o, "fso_boundary_residual", jac) == 0
        R, Rr = s.get("R"), o["R"]
        rerr = np.max(np.abs(R - Rr)) / np.max(np.abs(Rr))
        L, Lr = s.get("LHS").reshape(-1, 16), o["LHS"].reshape(-1, 16)
        off = np.ones(L.shape[0], bool)
        off[iau] = False
        np.testing.assert_array_equal(L[off], Lr[off])
        if not jac:
            assert not np.any(L[off]), "an iteration without a Jacobian update keeps only V/dt"
        derr = np.max(np.abs(L[iau] - Lr[iau])) / np.max(np.abs(Lr[iau]))
        np.testing.assert_array_equal(s.get("nQ"), o["nQ"])
        s.set("RHS", -R)
        o["RHS"][:] = -Rr
        o["dQ"][:] = 0.0
        s.linear_solve()
        U.oracle_call(o, "fso_linear_solve", 1)
        dQ, dQr = s.get("dQ"), o["dQ"]
        qerr = np.max(np.abs(dQ - dQr)) / np.max(np.abs(dQr))
        print("ts %d it %d jac %d: R err %.1e diag err %.1e dQ err %.1e" % (ts, it, jac, rerr, derr, qerr), flush=True)
        assert rerr <= 1e-12 and derr <= 1e-12 and qerr <= 1e-10
        # the next iteration starts from the same updated state on both sides
        nQ = o["nQ"].copy()
        nQ[:4 * nn] += dQr
        s.set("nQ", nQ)
        o["nQ"][:] = nQ
