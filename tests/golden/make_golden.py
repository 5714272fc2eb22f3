"""Writes the golden vectors under tests/golden/ from the reference build.

Runs oracle/_ref/ref_driver (the reference's own translation units, compiled
in place by oracle/Makefile from /root/reference) on seeded meshes written in
the reference grid format, and keeps the arrays the tests compare as
compressed .npz fixtures (0-based indices, reference 1-based arrays shifted).

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import fsutil as U  # noqa: E402

fsm = U.fsm

SETUP_INT = ["nn_map", "nnsn", "nsn", "esn", "edges", "csn", "subedges", "bedges", "ia", "iau", "ja", "c2n_tri"]
SETUP_DBL = ["x", "xn_subedges", "xm_subedges", "xg_subedges", "el_cent", "xn_bedges", "xg_bedges", "cv_area",
             "Moments", "edge_weight", "Q0"]


def params(**kw):
    base = dict(nsteps=6, iter=1, jufreq=1, CFL=20.0, ramp=10, gamma=1.4, alpha=1.0, mach=0.5, subits=10, method=1,
                order=1, foits=2, soits=0, limit=1, reorder=1)
    base.update(kw)
    return fsm.Params.make(**base)


def run_case(name, nx, ny, seed, p, nsteps, stencil, step_fields, dump_every=1, keep_steps=None, extra=None,
             setup=True, lhs_steps=(), airfoil=None, mixed=False):
    """airfoil = (first cell, outer radius, jitter): an O-mesh of nx wall x ny
    radial nodes instead of the bump channel; mixed: the bump channel with
    quadrilaterals (fsutil.bump_mixed)."""
    with tempfile.TemporaryDirectory() as tmp:
        if airfoil:
            mesh = fsm.Mesh.airfoil(nx, ny, airfoil[0], airfoil[1], airfoil[2], seed)
        elif mixed:
            mesh = U.bump_mixed(nx, ny, seed)
        else:
            mesh = fsm.Mesh.bump(nx, ny, 0.2, seed)
        U.write_case(tmp, mesh, p)
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([U.REF_DRIVER, tmp, out, str(nsteps), "1" if stencil else "0", str(dump_every)],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-3000:])
        d = U.read_dump(out)
    keep = {"params": np.array([nx, ny, seed, p.order, p.foits, p.soits, p.limit, nsteps], dtype=np.int64),
            "params_f": np.array([p.CFL, p.mach, p.alpha, p.ramp, p.subits], dtype=np.float64),
            "hdr": d["hdr"]}
    if airfoil:
        keep["airfoil"] = np.array(airfoil, dtype=np.float64)
    if mixed:
        keep["mixed"] = np.array([1], dtype=np.int64)
    if p.iter != 1 or p.jufreq != 1:
        keep["newton"] = np.array([p.iter, p.jufreq], dtype=np.int64)
    for k in (SETUP_INT + SETUP_DBL) if setup else ["Q0"]:
        keep["s_" + k] = d[k]
    if setup and "c2n_quad" in d:
        keep["s_c2n_quad"] = d["c2n_quad"]
    if stencil:
        for k in ("gnnsn1", "gnnsn", "gnsn"):
            keep["s_" + k] = d[k]
    steps = keep_steps if keep_steps is not None else range(1, nsteps + 1)
    for ts in steps:
        for f in step_fields:
            keep["%s_%d" % (f, ts)] = d["%s_%d" % (f, ts)]
        if ts in lhs_steps:
            keep["LHS_%d" % ts] = d["LHS_%d" % ts]
    if extra:
        for k in extra:
            keep[k] = d[k]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **keep)
    print("%s: %d arrays, %.0f KB" % (path, len(keep), os.path.getsize(path) / 1024))


def main():
    if not os.path.exists(U.REF_DRIVER):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    full = ["rms", "R", "dQ", "Q", "phi", "grad", "ev_con", "dt", "forces"]
    # tiny meshes: complete step histories through the first-order and linear phases
    run_case("bump17x9_o1", 17, 9, 7, params(order=1, foits=2), 6, False, full, lhs_steps=(1, 3, 6))
    run_case("bump17x9_o3", 17, 9, 7, params(order=3, foits=2, soits=3), 5, True, full, lhs_steps=(3, 5))
    # Venkatakrishnan limiter (limit=2) at the k=1 face points
    run_case("bump17x9_o1_venkat", 17, 9, 7, params(order=1, foits=2, limit=2, mach=0.675), 6, False, full,
             lhs_steps=(6,))
    # three Newton iterations per step, the Jacobian updated on the 1st and 3rd
    run_case("bump17x9_o1_newton", 17, 9, 7, params(order=1, foits=2, iter=3, jufreq=2), 5, False, full,
             lhs_steps=(5,))
    # O-mesh around the 12%-thick profile: closed wall, far field, seam of the
    # periodic direction; limited linear phase at an angle of attack
    run_case("airfoil32x9_o1", 32, 9, 4, params(order=1, foits=2, mach=0.5, alpha=2.0), 6, False, full,
             lhs_steps=(6,), airfoil=(2e-2, 8.0, 0.1))
    # mixed elements: the bump channel with a third of its cells kept as
    # quadrilaterals (4 subedges each), first-order and limited linear phases
    run_case("bumpmix17x9_o1", 17, 9, 7, params(order=1, foits=2), 6, False, full, lhs_steps=(6,), mixed=True)
    run_case("bumpmix17x9_o3", 17, 9, 7, params(order=3, foits=2, soits=3), 5, True, full, lhs_steps=(5,),
             mixed=True)
    # the CPU-runnable configuration: 97x33 bump channel, M=0.5, k=1, Barth limiter
    run_case("cfg1_97x33", 97, 33, 1, params(order=3, foits=20, soits=20, mach=0.5, alpha=0.0, CFL=100.0, ramp=30,
                                              subits=20), 2, True, ["rms", "R", "dQ", "Q", "forces"])
    # converged first-order-then-linear unlimited solve on the same mesh (1e-10 residual drop reached by step 174)
    run_case("cfg1_97x33_steady", 97, 33, 1, params(order=1, foits=20, limit=0, mach=0.5, alpha=0.0, CFL=100.0,
                                                     ramp=30, subits=20), 180, False, ["rms", "forces"], dump_every=20,
             keep_steps=[20, 40, 60, 80, 100, 120, 140, 160, 180], extra=["Qfinal"], setup=False)


if __name__ == "__main__":
    main()
