"""The reference-shaped C++ entry points (include/fs_reference_shim.hpp),
driven by tests/shim/shim_check (built by __graft_entry__.build()).

CPU: the host Roe_flux_centered of the shared per-face header equals the
oracle's restatement of flux.C:299-541 bit for bit; the curved_boundary_*
functions (curved_boundaries.h:25-33) return what the C-ABI returns.
GPU: a GRID in the reference layout advanced by fsref::solve, and one
reference call sequence through the loop-level wrappers, give the C-ABI
solver's arrays bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import fsutil as U

SHIM = os.path.join(U.ROOT, "tests", "shim", "shim_check")


def _need_shim():
    if not os.path.exists(SHIM):
        pytest.fail("tests/shim/shim_check is not built (run __graft_entry__.build())")


def test_host_roe_flux_bitexact_vs_oracle():
    _need_shim()
    rng = np.random.default_rng(11)
    n = 2000
    th = rng.uniform(0, 2 * np.pi, n)
    nrm = np.stack([np.cos(th), np.sin(th), rng.uniform(0.01, 1.0, n)], 1)

    def states():
        rho = rng.uniform(0.5, 2.0, n)
        u, v = rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n)
        p = rng.uniform(0.3, 2.0, n)
        return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], 1)

    qL, qR = states(), states()
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int64(n).tobytes())
            f.write(np.ascontiguousarray(np.concatenate([nrm, qL, qR], 1)).tobytes())
        r = subprocess.run([SHIM, "roe", fin, fout], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        out = np.fromfile(fout, dtype=np.float64).reshape(n, 4)
    ref = np.stack([U.roe_oracle(*nrm[i], 1.4, qL[i], qR[i]) for i in range(n)])
    assert np.array_equal(out, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("order,limit", [(1, 1), (3, 1), (1, 2)])
def test_reference_shaped_solve_and_sequence(order, limit):
    _need_shim()
    print("shim_check solve 33 17 5 order=%d limit=%d" % (order, limit), flush=True)
    r = subprocess.run([SHIM, "solve", "33", "17", "5", str(order), str(limit), "6"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "ALL IDENTICAL" in r.stdout


REF_DRIVER_GPU = os.path.join(U.ORACLE_DIR, "_ref", "ref_driver_gpu")


@pytest.mark.gpu
def test_reference_driver_with_device_solve():
    """The reference's own driver loop with its solve() replaced by the device
    path (oracle/ref_gpu_solve.cpp, the reference-side binding) converges to
    the same steady state as the unmodified reference build (1e-10 in the
    density L2 norm).  The iterate paths differ (multicolor against
    row-sequential Gauss-Seidel); the fixed point does not."""
    for exe in (U.REF_DRIVER, REF_DRIVER_GPU):
        if not os.path.exists(exe):
            pytest.skip("reference build harness not built here (needs the reference sources at build time)")
    fsm = U.fsm
    p = fsm.Params.make(nsteps=260, iter=1, jufreq=1, CFL=100.0, ramp=30, gamma=1.4, alpha=0.0, mach=0.5,
                        subits=20, method=1, order=1, foits=10, soits=0, limit=0, reorder=1)
    n = 260
    with tempfile.TemporaryDirectory() as tmp:
        m = fsm.Mesh.bump(49, 17, 0.2, 1)
        U.write_case(tmp, m, p)
        out = {}
        for exe, tag in ((U.REF_DRIVER, "cpu"), (REF_DRIVER_GPU, "gpu")):
            print("running", os.path.basename(exe), flush=True)
            r = subprocess.run([exe, tmp, os.path.join(tmp, tag + ".bin"), str(n), "0", str(n)], capture_output=True,
                               text=True, timeout=300, cwd=tmp)
            assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
            out[tag] = U.read_dump(os.path.join(tmp, tag + ".bin"))
    nn = int(out["cpu"]["hdr"][0]) if "hdr" in out["cpu"] else m.info()["nn"]
    qc, qg = out["cpu"]["Q_%d" % n], out["gpu"]["Q_%d" % n]
    rc, rg = qc[4:4 + 4 * nn:4], qg[4:4 + 4 * nn:4]
    rel = np.linalg.norm(rg - rc) / np.linalg.norm(rc)
    print("reference driver, CPU solve vs device solve after %d steps: density L2 rel diff %.3e; |dQ| %.3e vs %.3e"
          % (n, rel, out["cpu"]["rms_%d" % n][0], out["gpu"]["rms_%d" % n][0]))
    assert out["gpu"]["rms_%d" % n][0] < 1e-9
    assert rel <= 1e-10
    # Compute_Body_Forces after the last step: the reference's own function on
    # the CPU build, the device one (the binding replaces it too) on the other
    fc, fg = out["cpu"]["forces_%d" % n], out["gpu"]["forces_%d" % n]
    print("body forces CPU %s device %s" % (fc, fg))
    np.testing.assert_allclose(fg, fc, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("bc,xL,xR", [(17, (0.1, 0.04683), (0.2, 0.05738)), (19, (0.6, -0.0456), (0.5, -0.0529)),
                                      (11, (1.0, 0.0), (0.9238795325112867, 0.3826834323650898)),
                                      (22, (300.0, 0.0), (0.0, 300.0))])
def test_curved_boundary_functions(bc, xL, xR):
    _need_shim()
    out = subprocess.run([SHIM, "curved", str(bc), repr(xL[0]), repr(xL[1]), repr(xR[0]), repr(xR[1])],
                         capture_output=True, text=True, check=True).stdout.split()
    v = [float(t) for t in out]
    assert tuple(v[0:2]) == U.fsm.curved_midpoint(bc, xL, xR)
    gp = U.fsm.curved_gauss_points(bc, xL, xR)
    assert [tuple(v[2 + 2 * k:4 + 2 * k]) for k in range(3)] == gp
    assert tuple(v[8:10]) == U.fsm.curved_normal(bc, xL)
    assert v[10] == U.fsm.curved_arclength(bc, xL, xR)
