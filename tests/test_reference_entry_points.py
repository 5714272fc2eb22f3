"""The reference-shaped entry points of include/fs_reference_shim.hpp against
the reference's own compiled functions (oracle/_ref/ref_pointwise, built from
the reference's translation units by oracle/Makefile), bit for bit, over
every face and boundary edge of seeded meshes:

  * host: Reconstruct, Reconstruct_Gauss and Reconstruct_Gauss_Boundary in
    the first-order, linear and k=2 phases (the face point and the three
    Gauss points of every subedge and boundary edge, from arbitrary limiter
    values, derivatives and point values); Get_Boundary_Value at the node
    and the Gauss points of every boundary edge; compute_numerical_jacobian_roe,
    compute_approximate_jacobian_roe (at unequal reconstructed states),
    compute_boundary_approximate_jacobian_roe and
    compute_boundary_numerical_jacobian_roe accumulated over every face into
    SYS_MEM's blocks; Build_CRS; Compute_Edge_Weights;
  * device: Set_Boundary_Values (once, and twice to exercise the ghosts'
    memory) and cv_calc_eigenvalue_contribution as device passes;
  * the oracle's analytic Jacobian against the reference's
    compute_approximate_jacobian_roe (the harness's jac mode writes the
    reference's blocks and the state they came from).

Meshes: the jittered bump channel, the same channel with a third of its cells
quadrilaterals, and the O-mesh around the 12%-thick profile (polyline wall).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import fsutil as U

fsm = U.fsm
HARNESS = os.path.join(U.ORACLE_DIR, "_ref", "ref_pointwise")


def _params():
    return fsm.Params.make(nsteps=6, iter=1, jufreq=1, CFL=20.0, ramp=10, gamma=1.4, alpha=1.5, mach=0.55, subits=10,
                           method=1, order=3, foits=2, soits=2, limit=1, reorder=1)


def _mesh(kind):
    if kind == "bump":
        return fsm.Mesh.bump(17, 9, 0.2, 7)
    if kind == "mixed":
        return U.bump_mixed(17, 9, 7)
    return fsm.Mesh.airfoil(32, 9, 2e-2, 8.0, 0.1, 4)


def _run(kind, mode):
    if not os.path.exists(HARNESS):
        pytest.skip("oracle/_ref/ref_pointwise is not built (make -C oracle ref needs the reference sources)")
    with tempfile.TemporaryDirectory() as tmp:
        U.write_case(tmp, _mesh(kind), _params())
        print("ref_pointwise %s %s" % (kind, mode), flush=True)
        r = subprocess.run([HARNESS, tmp, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    checks = [ln for ln in r.stdout.splitlines() if ln.startswith("check ")]
    assert r.returncode == 0 and "ALL BIT-EXACT" in r.stdout
    return checks


@pytest.mark.parametrize("kind", ["bump", "mixed", "airfoil"])
def test_host_entry_points_bitexact(kind):
    checks = _run(kind, "host")
    assert len(checks) == 10


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bump", "airfoil"])
def test_device_entry_points_bitexact(kind):
    checks = _run(kind, "device")
    assert len(checks) == 3


@pytest.mark.parametrize("kind", ["bump", "mixed", "airfoil"])
def test_oracle_analytic_jacobian_matches_reference(kind):
    """The oracle's analytic Jacobian (its restatement of jacobian.C:301-658,
    which the mixed path's FP32 blocks are checked against) equals the
    reference's own compiled compute_approximate_jacobian_roe, accumulated
    over every face of the mesh at the linear phase's reconstructed face
    states: a perturbed state and arbitrary limiter values make the two face
    states of every face unequal, so the Roe average's weighting, the flux
    Jacobians of both sides and the signs of the four block updates all
    enter.  Bit for bit."""
    if not os.path.exists(HARNESS):
        pytest.skip("oracle/_ref/ref_pointwise is not built (make -C oracle ref needs the reference sources)")
    p = _params()
    m = _mesh(kind)
    with tempfile.TemporaryDirectory() as tmp:
        U.write_case(tmp, m, p)
        r = subprocess.run([HARNESS, tmp, "jac"], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        nn, ng, nnz, ts = (int(v) for v in open(os.path.join(tmp, "jac_meta.txt")).read().split())
        nq = np.fromfile(os.path.join(tmp, "jac_nQ.bin"), dtype=np.float64)
        phi = np.fromfile(os.path.join(tmp, "jac_phi.bin"), dtype=np.float64)
        lhs = np.fromfile(os.path.join(tmp, "jac_lhs.bin"), dtype=np.float64)
    m.prepare(p)
    o = U.Oracle(m, p)
    assert (o.info["nn"], o.info["ng"], o.info["nnz"]) == (nn, ng, nnz)
    o["nQ"][:] = nq[4:]
    o["phi"][:] = phi[8:]
    U.oracle_call(o, "fso_gradient")
    o["LHS"][:] = 0.0
    L = U.oracle_lib()
    L.fso_set_analytic_jacobian(1)
    try:
        U.oracle_call(o, "fso_interior_residual", 1, ts)
    finally:
        L.fso_set_analytic_jacobian(0)
    # the face states really differ (else the Roe average would equal both)
    be = m.array("edges").reshape(-1, 2)
    assert np.min(np.abs(nq[4:][4 * be[:, 0]] - nq[4:][4 * be[:, 1]])) > 0.0
    nz = np.count_nonzero(lhs)
    mism = int(np.count_nonzero(o["LHS"] != lhs))
    err = np.max(np.abs(o["LHS"] - lhs)) / np.max(np.abs(lhs))
    print("%s: %d nonzero block entries, %d differ in any bit, max rel diff %.2e" % (kind, nz, mism, err), flush=True)
    assert nz > 0
    assert mism == 0
