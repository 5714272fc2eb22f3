"""Benchmark: edge-flux evaluations per second of the residual-and-implicit-
update path on one MI355X.

Workload (default): bump channel 1025x257 nodes (524,288 triangles, seeded
+-0.2 jitter), M=0.675, k=2 (order 3) with the Barth limiter, FP64, the
reference's implicit pseudo-time step (finite-difference Jacobian, 20
symmetric Gauss-Seidel sweeps).  Warmup steps walk through the first-order and
linear phases into the k=2 phase; the timed steps are k=2-phase steps.

One step = one pseudo-time step (one Newton iteration): residual + Jacobian +
linear sweeps + update over the whole mesh.  Edge-flux evals per step =
subedges + boundary edges (the residual's Roe evaluations; the 8 extra
Jacobian evaluations per subedge are reported separately).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_edge_kernel(info, recon):
    """Compulsory HBM bytes of one edge flux+Jacobian launch (FP64)."""
    ne, nsub, nn = info["ne"], info["nsub"], info["nn"]
    b = ne * (8 + 8 + 8)            # (L,R), subedge ids, the two block slots
    b += nsub * 24                  # (nx, ny, len)
    b += nsub * 32                  # per-subedge flux written
    b += ne * 2 * 128               # two off-diagonal 4x4 FP64 blocks written
    b += nn * 32                    # node states read once
    if recon == 1:
        b += nsub * 16 + nn * (64 + 8 + 16)   # face point; grad, phi, x per node
    return b


def gs_schedule(ncolor, subits):
    """Color visiting order of the device sweeps (turnaround repeats dropped)."""
    out, last = [], -1
    for it in range(subits):
        for cc in range(ncolor):
            c = cc if it % 2 == 0 else ncolor - 1 - cc
            if c != last:
                out.append(c)
                last = c
    return out


def algorithmic_bytes_gauss_seidel(mesh, subits, word):
    """Compulsory HBM bytes of one linear solve: the diagonal factorisation
    pass plus every color launch of the sweep schedule.  word = 8 (FP64) or 4.
    Per launch of color c: its rows' off-diagonal blocks (16 words) and column
    indices, the rows' factored diagonals (16 words), RHS and dQ (4 words each,
    dQ read and written), and the neighbours' dQ once (4 words per node of the
    other colors)."""
    import numpy as np
    info = mesh.info()
    nn = info["nn"]
    ia = mesh.array("ia")
    deg = np.diff(ia) - 1
    co, rows = mesh.array("color_off"), mesh.array("color_rows")
    per_color = []
    for c in range(info["ncolor"]):
        r = rows[co[c]:co[c + 1]]
        nr, nb = len(r), int(deg[r].sum())
        per_color.append(nb * (16 * word + 4) + nr * (16 * word + 12 * word + 12) + (nn - nr) * 4 * word)
    sched = gs_schedule(info["ncolor"], subits)
    total = sum(per_color[c] for c in sched)
    return total, len(sched)


def measured_traffic(kernel):
    """HBM bytes per dispatch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/pmc_traffic_bump1025x257.md gives the commands)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_bump1025x257.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--nx", type=int, default=1025)
    ap.add_argument("--ny", type=int, default=257)
    ap.add_argument("--cpu-steps", type=int, default=2, help="k=2-phase steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline(U, fsm, p, nx, ny, steps):
    """The oracle (single-threaded C restatement) on the same workload: the
    untimed steps up to the k=2 phase, then `steps` timed k=2-phase steps."""
    mesh = fsm.Mesh.bump(nx, ny, 0.2, 2)
    mesh.prepare(p)
    info = mesh.info()
    o = U.Oracle(mesh, p)
    o.init_freestream()
    first = p.foits + p.soits + 1
    for ts in range(1, first):
        o.solve(ts, multicolor=0)
    t0 = time.perf_counter()
    for ts in range(first, first + steps):
        o.solve(ts, multicolor=0)
    dt = time.perf_counter() - t0
    evals = (info["nsub"] + info["nbe"]) * steps
    return {"value": evals / dt, "unit": "edge-flux evals/s", "cores": 1, "kind": "port",
            "sample": "%d k=2-phase pseudo-time steps of the same %dx%d bump channel (%.1f s)" % (steps, nx, ny, dt)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import fsutil as U

    fsm = U.fsm
    p = fsm.Params.make(nsteps=a.warmup + a.steps, iter=1, jufreq=1, CFL=50.0, ramp=20, gamma=1.4, alpha=0.0,
                        mach=0.675, subits=20, method=1, order=3, foits=2, soits=2, limit=1, reorder=1)
    t0 = time.perf_counter()
    mesh = fsm.Mesh.bump(a.nx, a.ny, 0.2, 2)
    mesh.prepare(p)
    t_setup = time.perf_counter() - t0
    info = mesh.info()
    solver = fsm.Solver(mesh, p, fsm.FS_PREC_F64)
    solver.init_freestream()
    warm = max(a.warmup, p.foits + p.soits)
    for ts in range(1, warm + 1):
        solver.solve(ts)
    tim = solver.time_steps(warm + 1, a.steps)
    ms = tim["step_ms"]
    evals = info["nsub"] + info["nbe"]
    value = evals / (ms * 1e-3)
    abytes = algorithmic_bytes_edge_kernel(info, recon=0)
    edge_gbs = abytes / (tim["edge_ms"] * 1e-3) / 1e9
    gs_bytes, gs_launches = algorithmic_bytes_gauss_seidel(mesh, p.subits, 8)
    achieved = gs_bytes / (tim["gs_ms"] * 1e-3) / 1e9
    line = {
        "metric": "edge-flux evals/s", "value": value, "unit": "edge-flux evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": warm, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "bump channel %dx%d, M=0.675, k=2 (order 3), Barth limiter, FP64 implicit step"
                               % (a.nx, a.ny), "nodes": info["nn"], "subedges": info["nsub"],
                   "bedges": info["nbe"], "lhs_blocks": info["nnz"], "precision": "fp64", "subits": p.subits},
        "roofline": {"bound": "hbm", "kernel": "k_gs_color (multicolor block Gauss-Seidel sweep, FP64)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (None if (a.nx, a.ny) != (1025, 257) else measured_traffic("fsk::k_gs_color")),
                     "algorithmic_bytes": gs_bytes / gs_launches, "launches_per_step": gs_launches,
                     "avg_ms": tim["gs_ms"] / gs_launches},
        "edge_kernel_roofline": {"kernel": "k_edge_flux (Roe flux + FD Jacobian blocks, FP64 VALU)",
                                 "achieved_GBs": edge_gbs, "algorithmic_bytes": abytes, "avg_ms": tim["edge_ms"]},
        "kernel_ms": {"edge_flux": tim["edge_ms"], "node_assembly": tim["node_ms"], "gauss_seidel": tim["gs_ms"],
                      "diag_lu": tim["lu_ms"], "other": tim["other_ms"]},
        "jacobian_flux_evals_per_step": 8 * info["nsub"],
        "setup_s": t_setup,
    }
    # the mixed-precision path on the same workload and the same timed steps
    sm = fsm.Solver(mesh, p, fsm.FS_PREC_MIXED)
    sm.init_freestream()
    for ts in range(1, warm + 1):
        sm.solve(ts)
    tm = sm.time_steps(warm + 1, a.steps)
    sm.close()
    gs32_bytes, _ = algorithmic_bytes_gauss_seidel(mesh, p.subits, 4)
    line["mixed"] = {"value": evals / (tm["step_ms"] * 1e-3), "ms_per_step": tm["step_ms"],
                     "gauss_seidel_f32_GBs": gs32_bytes / (tm["gs_ms"] * 1e-3) / 1e9,
                     "speedup_vs_fp64": ms / tm["step_ms"], "dtype": "f64 residual/norms/update, f32 Jacobian+sweeps",
                     "kernel_ms": {"edge_flux_f64": tm["edge_ms"], "node_assembly": tm["node_ms"],
                                   "gauss_seidel_f32": tm["gs_ms"], "diag_lu_f32": tm["lu_ms"],
                                   "other": tm["other_ms"]}}
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(U, fsm, p, a.nx, a.ny, a.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
