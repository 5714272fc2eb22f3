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
    memory) and cv_calc_eigenvalue_contribution as device passes.

Meshes: the jittered bump channel, the same channel with a third of its cells
quadrilaterals, and the O-mesh around the 12%-thick profile (polyline wall).
"""
import os
import subprocess
import tempfile

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
