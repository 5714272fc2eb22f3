"""Seeded synthetic workloads: the batch of independent steady solves and the
serving request stream.

A case is drawn from its own seed with numpy's PCG64 generator; the mesh
itself comes from the library's seeded generators (SplitMix64-seeded
xoshiro256** jitter), so a (kind, size, seed) triple always names the same
grid.

  * batch case: bump channel or airfoil O-mesh with a triangle count
    log-uniform in [5k, 200k], Mach ~ U[0.3, 0.8], angle of attack
    ~ U[-2, 4] degrees, reconstruction k in {1, 2};
  * serving request: the same with triangles log-uniform in [10k, 2M] and
    Mach ~ U[0.3, 0.85].

Triangle counts: a bump channel of nx*ny nodes has 2(nx-1)(ny-1) triangles
(nx ~ 3 ny for the 3:1 channel); an O-mesh of nwall*nrad nodes has
2 nwall (nrad-1) (nwall ~ 4 nrad).
"""
import concurrent.futures as cf
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import flowsolver as fsm  # noqa: E402


def bump_dims(tri):
    ny = max(3, int(round(math.sqrt(tri / 6.0))) + 1)
    nx = max(3, int(round(tri / (2.0 * (ny - 1)))) + 1)
    return nx, ny


def omesh_dims(tri):
    nrad = max(4, int(round(math.sqrt(tri / 8.0))) + 1)
    nwall = max(16, int(round(tri / (2.0 * (nrad - 1)))))
    return nwall, nrad


def draw_case(seed, tri_lo, tri_hi, mach_lo, mach_hi):
    """One solve: mesh description and PARAMS fields, from its seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kind = "bump" if rng.random() < 0.5 else "omesh"
    tri = int(round(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))))
    mach = float(rng.uniform(mach_lo, mach_hi))
    alpha = float(rng.uniform(-2.0, 4.0))
    k = 1 if rng.random() < 0.5 else 2
    dims = bump_dims(tri) if kind == "bump" else omesh_dims(tri)
    return {"seed": int(seed), "kind": kind, "dims": dims, "mach": mach, "alpha": alpha, "k": k}


def case_params(case, **over):
    """PARAMS for a case: implicit ramped CFL, first-order start-up, Barth
    limiter, the reference's 20 sweeps; k=1 -> order 1, k=2 -> order 3."""
    kw = dict(nsteps=0, iter=1, jufreq=1, CFL=50.0, ramp=20, gamma=1.4, alpha=case["alpha"], mach=case["mach"],
              subits=20, method=1, order=1 if case["k"] == 1 else 3, foits=10, soits=10 if case["k"] == 2 else 0,
              limit=1, reorder=1)
    kw.update(over)
    return fsm.Params.make(**kw)


def make_mesh(case):
    if case["kind"] == "bump":
        nx, ny = case["dims"]
        return fsm.Mesh.bump(nx, ny, 0.2, case["seed"])
    nwall, nrad = case["dims"]
    return fsm.Mesh.airfoil(nwall, nrad, 1e-3, 30.0, 0.1, case["seed"])


def build_cases(cases, params, workers=8):
    """Generate and prepare the meshes (host preprocessing runs in parallel:
    the library releases the interpreter lock inside its calls)."""
    def one(i):
        m = make_mesh(cases[i])
        m.prepare(params[i])
        return m
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(len(cases))))


def batch_cases(n=256, seed0=1000, tri_lo=5000, tri_hi=200000):
    return [draw_case(seed0 + i, tri_lo, tri_hi, 0.3, 0.8) for i in range(n)]


def serving_cases(n, seed0=5000, tri_lo=10000, tri_hi=2000000):
    return [draw_case(seed0 + i, tri_lo, tri_hi, 0.3, 0.85) for i in range(n)]
