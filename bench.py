"""Benchmark: edge-flux evaluations per second of the residual-and-implicit-
update path on one MI355X, mixed precision against the plain FP64 path, and
steady-solve latency of the serving loop.

Headline workload: 256 independent steady solves packed into one device
engine (bump channels and airfoil O-meshes, 5k-200k triangles log-uniform,
M ~ U[0.3, 0.8], alpha ~ U[-2, 4] deg, k in {1, 2}; seeded), every kernel of a
pseudo-time step one launch over all instances, mixed precision (FP64
residual, norms and update; FP32 analytic Jacobian and Gauss-Seidel sweeps).
Batch and serving solves run with recovery from non-physical states, an
opt-in control the reference lacks (see WORKLOAD_CONTROLS).
Warm-up steps walk every instance into its final phase (linear reconstruction
for k=1, the k=2 phase for k=2); the timed steps are final-phase steps, timed
with events on the engine's stream, which also give each kernel's time.

One step = one pseudo-time step (one Newton iteration) of every instance:
residual + Jacobian + 40 symmetric Gauss-Seidel sweeps + update.  Edge-flux
evals per step = subedges + boundary edges of all instances (the residual's
Roe evaluations).

Also measured (extra objects of the same line):
  * "fp64": the same batch and steps on the plain FP64 path (the reference's
    finite-difference Jacobian and FP64 sweeps);
  * "steady": every batch solve run to a 1e-10 drop under per-instance
    convergence masks, both precisions, converged states compared;
  * "plain_fp64_sequential": the same solves one at a time on the FP64 path;
  * "config2": the 1025x257 bump channel at M=0.675, k=2, FP64 and mixed;
  * "config3": the 600k-node transonic O-mesh, FP64 and mixed to 1e-10;
  * "serving": a seeded Poisson stream of steady-solve requests (10k-2M
    triangles) through the C++ request loop, p50/p99 time-to-converged with
    unconverged requests counted as censored, mixed and FP64;
  * CPU baselines: the oracle on the host cores (one batch solve per core),
    the oracle on config 2 (1 core), and the reference build itself.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "flow-solver_amd"))

HBM_PEAK_GBS = 8000.0
_T0 = time.time()


def progress(msg):
    """A progress line on stderr with the seconds since the bench started."""
    print("[%6.1f s] %s" % (time.time() - _T0, msg), file=sys.stderr, flush=True)


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


def algorithmic_bytes_resid_jac32(info, recon, fused=0, block_word=4):
    """Compulsory HBM bytes of one launch of the mixed path's node pass
    k_resid_jac32 (FP64 residual, ev_con/dt, FP32 Jacobian blocks):
    per node the state read (32 B), R, RHS, RHS32, ev_con and dt written
    (32 + 32 + 16 + 8 + 8 B), cv_area and the diagonal slot read (8 + 4 B),
    the FP32 diagonal block written (64 B) and the activity byte (1 B), and
    when the pass forms the pre-multiplied sweep operands itself (fused) the
    zeroed dQ (16 B); per subedge its normal and length (24 B); per edge its
    two node ids and two block slots (16 B) and its two off-diagonal blocks
    written (16 block words each: FP32, or FP16 operands); per boundary edge
    its normal, flux and block read (24 + 32 + 128 B).  With linear
    reconstruction (recon 1) also the gradient, limiter value and
    coordinates per node (64 + 8 + 16 B) and the face point per subedge
    (16 B)."""
    b = info["nn"] * (205 + 16 * fused) + info["nsub"] * 24 + info["ne"] * (16 + 32 * block_word) + info["nbe"] * 184
    if recon == 1:
        b += info["nn"] * 88 + info["nsub"] * 16
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


def gs_row_bytes(word, scaled=0):
    """Per-row bytes of one sweep launch besides the off-diagonal blocks.
    FP64 (word 8): the factored diagonal (16 words), RHS, dQ read and dQ
    written (4 words each) and 12 bytes of row indices (node, degree,
    instance).  Mixed path (word 4, vectors by color position): the diagonal
    inverse (16 words) unless the operands are pre-multiplied by it
    (scaled), the right-hand side and dQ written (4 words each), the degree
    and the activity byte (5 bytes)."""
    if word == 8:
        return 16 * word + 12 * word + 12
    return (8 * word + 5) if scaled else (24 * word + 5)


def algorithmic_bytes_gauss_seidel(mesh, subits, word, scaled=0, block_word=None):
    """Compulsory HBM bytes of one linear solve's sweeps: every color launch
    of the sweep schedule.  word = 8 (FP64) or 4; block_word (default word)
    = 2 for FP16 operands.  Per launch of color c: its rows' off-diagonal
    blocks (16 block words) and column indices, the per-row operands
    (gs_row_bytes), and the neighbours' dQ once (4 words per node of the
    other colors)."""
    block_word = block_word or word
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
        per_color.append(nb * (16 * block_word + 4) + nr * gs_row_bytes(word, scaled) + (nn - nr) * 4 * word)
    sched = gs_schedule(info["ncolor"], subits)
    total = sum(per_color[c] for c in sched)
    return total, len(sched)


def algorithmic_bytes_gs_batch(meshes, subits, word, scaled=0, block_word=None):
    """The same for the sweeps of a packed batch, launch by launch as the
    engine issues them: one launch per global color of the shared schedule
    (over the largest color count), each over that color's rows of every
    instance; a backward launch of color c skips the instances whose last
    color is c (they repeat no turnaround), and an empty launch is not
    issued.  Returns (total bytes per linear solve, launches)."""
    import numpy as np
    block_word = block_word or word
    per = []
    for m in meshes:
        info = m.info()
        deg = np.diff(m.array("ia")) - 1
        co, rows = m.array("color_off"), m.array("color_rows")
        nn = info["nn"]
        cols = []
        for c in range(info["ncolor"]):
            r = rows[co[c]:co[c + 1]]
            nr, nb = len(r), int(deg[r].sum())
            cols.append(nb * (16 * block_word + 4) + nr * gs_row_bytes(word, scaled) + (nn - nr) * 4 * word)
        per.append((info["ncolor"], cols))
    ncmax = max(nc for nc, _ in per)
    total, launches, last = 0, 0, -1
    for it in range(subits):
        for cc in range(ncmax):
            c = cc if it % 2 == 0 else ncmax - 1 - cc
            if last == c:
                continue
            last = c
            bwd = it % 2 == 1
            b = sum(cols[c] for nc, cols in per if c < nc and not (bwd and nc - 1 == c))
            if b == 0:
                continue
            total += b
            launches += 1
    return total, launches


def measured_traffic(kernel, name="pmc_traffic_bump1025x257.json"):
    """HBM bytes per dispatch from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary under profiles/ (its .md twin gives the commands)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline batch")
    ap.add_argument("--warmup", type=int, default=20, help="warm-up steps (at least the final-phase start)")
    ap.add_argument("--config2-steps", type=int, default=20)
    ap.add_argument("--config2-warmup", type=int, default=6)
    ap.add_argument("--nx", type=int, default=1025)
    ap.add_argument("--ny", type=int, default=257)
    ap.add_argument("--cpu-steps", type=int, default=2, help="k=2-phase steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="instances of the headline batch")
    ap.add_argument("--batch-max-steps", type=int, default=2500, help="steady batch solve limit (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=55.0, help="length of the request stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="offered requests per second")
    ap.add_argument("--serve-drain", type=float, default=100.0,
                    help="seconds after the last arrival before unfinished requests are cancelled (counted as "
                         "never converged)")
    ap.add_argument("--serve-max-steps", type=int, default=2000)
    ap.add_argument("--serve-precisions", default="mixed,fp64", help="streams to run, in this order")
    ap.add_argument("--no-batch-plain", action="store_true", help="skip the one-solve-at-a-time FP64 baseline")
    ap.add_argument("--batch-plain-steady", action="store_true",
                    help="also run every batch solve alone to its stopping test in FP64 (about 90 s more)")
    ap.add_argument("--config3-max-steps", type=int, default=3000)
    ap.add_argument("--config3-steps", type=int, default=0,
                    help="instead of steady solves, this many k=2-phase steps per precision (polyline wall only, "
                         "one kernel launch at a time; for kernel traces)")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the extras named in --sections (profiles of one section)")
    ap.add_argument("--no-config3-curved", action="store_true", help="skip the analytic-wall variant of config 3")
    ap.add_argument("--sections", default="fp64,mixed16,steady,plain,config2,config3,serving,cpu",
                    help="extras beside the headline (a profiled run picks few): fp64, mixed16, steady, plain, "
                         "config2, config3, serving, cpu")
    return ap.parse_args()


# solve controls of the batch and serving workloads: implicit ramped CFL, 40
# sweeps, first-order start-up; k=1 unlimited (the limited k=1 iteration,
# Barth's or Venkatakrishnan's, cycles near a 1e-1 drop), k=2 with the Barth
# limiter (its k=2-phase residual is first order, so the limiter does not
# move the solution).  Recovery, an opt-in control the reference does not
# have: a solve that goes non-physical returns to its checkpoint with half
# its CFL target and a restarted ramp (at most RECOVERY[0] times,
# checkpointed every RECOVERY[1] steps); the first failure of a k=1 solve
# instead restarts its linear phase with the Barth limiter, frozen after
# RECOVERY[2] limited steps, and the next one restarts it at first order
# (RECOVERY[3]) at the solve's own CFL target; unlimited k=1 goes
# non-physical on transonic channels that choke, at any CFL target.  A solve
# whose residual has not halved in STALL_STEPS_PER_SQRT_NODE * sqrt(nodes)
# steps moves on to the next fallback.  A solve a fallback acted on converges to the
# steady state of another discretization (the frozen-limiter one depends on
# its iterate path): mixed and FP64 results are compared apart from it.
WORKLOAD_CONTROLS = dict(CFL=200.0, ramp=30, subits=40)
RECOVERY = (4, 10, 50, True)
STALL_STEPS_PER_SQRT_NODE = 2.0
# serving: a request's step limit grows with the square root of its node
# count (the steps to converge do, measured on the batch)
SERVE_STEPS_PER_SQRT_NODE = 8.0
# the 600k-node transonic O-mesh: at CFL 200 and 40 sweeps the FP64 path is
# still at a 4e-8 drop after 3000 steps and the mixed path (analytic
# approximate Jacobian) reaches a non-physical state.  CFL 1000 brings both
# to a 1e-10 drop on both walls: in about 1460 steps with 80 sweeps, in 830
# steps at about the same wall time with 160 (measured)
CONFIG3_CONTROLS = dict(CFL=1000.0, ramp=50, subits=160)


def workload_params(W, cases):
    return [W.case_params(c, limit=(0 if c["k"] == 1 else 1), **WORKLOAD_CONTROLS) for c in cases]


def workload_controls():
    return dict(WORKLOAD_CONTROLS, tol=1e-10, limiter="k=1 none, k=2 Barth", max_recoveries=RECOVERY[0],
                checkpoint_every=RECOVERY[1], limiter_fallback_freeze=RECOVERY[2], first_order_fallback=RECOVERY[3],
                stall_steps_per_sqrt_node=STALL_STEPS_PER_SQRT_NODE)


def apply_controls(b):
    """The workloads' opt-in recovery policy on a batch."""
    b.set_recovery(*RECOVERY)
    b.set_stall(STALL_STEPS_PER_SQRT_NODE)


def outcome_groups(fsm, cases, status):
    """Solve outcomes by mesh kind and reconstruction order."""
    groups = {}
    for c, st in zip(cases, status):
        g = groups.setdefault("%s_k%d" % (c["kind"], c["k"]), {"converged": 0, "max_steps": 0, "failed": 0,
                                                                "cancelled": 0})
        g[{fsm.INST_CONVERGED: "converged", fsm.INST_MAX_STEPS: "max_steps",
           fsm.INST_CANCELLED: "cancelled"}.get(int(st), "failed")] += 1
    return groups


def build_batch(a):
    """The 256-solve workload: seeded cases, PARAMS, prepared meshes."""
    import workloads as W
    cases = W.batch_cases(a.batch)
    params = workload_params(W, cases)
    t0 = time.perf_counter()
    meshes = W.build_cases(cases, params, workers=16)
    setup = time.perf_counter() - t0
    infos = [m.info() for m in meshes]
    desc = {"instances": a.batch, "triangles": sum(i["ntri"] for i in infos), "nodes": sum(i["nn"] for i in infos),
            "subedges": sum(i["nsub"] for i in infos), "bedges": sum(i["nbe"] for i in infos),
            "edges": sum(i["ne"] for i in infos), "k2_instances": sum(1 for c in cases if c["k"] == 2),
            "setup_s": setup, "controls": workload_controls()}
    return cases, params, meshes, infos, desc


def batch_throughput(a, fsm, meshes, params, infos, prec):
    """One engine over every instance: freestream start, one step, launch
    shapes tuned on the engine's data (results unchanged), hipGraph warm-up
    steps into every instance's final phase, `a.steps` hipGraph steps timed on
    the host clock, then `a.steps` steps launched one kernel at a time with
    events on the engine's stream around each kernel group (the headline
    region: step time and per-kernel time come from the same steps)."""
    warm = max(a.warmup, max(p.foits + p.soits for p in params))
    b = fsm.Batch(meshes, params, prec)
    b.use_graph(True)
    b.init_freestream()
    apply_controls(b)
    b.step()
    b.autotune()
    for _ in range(warm - 1):
        b.step()
    tg = b.time_steps(a.steps)
    b.use_graph(False)
    t = b.time_steps(a.steps)
    mixed_path = prec != fsm.FS_PREC_F64
    tuning = b.tuning()
    out = {"value": t["flux_evals"] / (t["step_ms"] * 1e-3), "unit": "edge-flux evals/s", "batch": len(meshes),
           "ms_per_step": t["step_ms"], "flux_evals_per_step": t["flux_evals"], "warmup": warm,
           "graph_ms_per_step": tg["step_ms"], "graph_value": tg["flux_evals"] / (tg["step_ms"] * 1e-3),
           "device_bytes": b.info()["device_bytes"], "tuning": tuning,
           "kernel_ms": {"boundary" if mixed_path else "edge_flux_fd_jacobian": t["edge_ms"],
                         "residual_jacobian_node_pass" if mixed_path else "node_assembly": t["node_ms"],
                         "gauss_seidel": t["gs_ms"], "diag_factor": t["lu_ms"], "other": t["other_ms"]}}
    b.close()
    word = 4 if mixed_path else 8
    block_word = 2 if prec == fsm.FS_PREC_MIXED16 else word
    scaled = tuning.get("gs32_scaled", 0) if mixed_path else 0
    gs_bytes, launches = algorithmic_bytes_gs_batch(meshes, params[0].subits, word, scaled, block_word)
    avg = t["gs_ms"] / launches
    out["gauss_seidel"] = {"kernel": sweep_kernel_name(prec, tuning), "launches_per_step": launches,
                           "algorithmic_bytes_per_launch": gs_bytes / launches,
                           "avg_ms": avg, "achieved_GBs": gs_bytes / launches / (avg * 1e-3) / 1e9}
    if mixed_path:
        fused = 1 if prec == fsm.FS_PREC_MIXED16 or (scaled and tuning.get("node_fused", 0)) else 0
        rb = sum(algorithmic_bytes_resid_jac32(i, 0 if p.order == 3 else 1, fused, block_word)
                 for i, p in zip(infos, params))
        kname = "fsk::k_resid_jac32<%d>" % (2 if prec == fsm.FS_PREC_MIXED16 else fused)
        out["residual_kernel"] = {
            "kernel": "%s (FP64 residual, ev_con/dt, FP32 Jacobian blocks and right-hand sides, one node pass%s)"
                      % (kname, "; forms the pre-multiplied sweep operands" if fused else ""),
            "avg_ms": t["node_ms"], "algorithmic_bytes": rb,
            "achieved_GBs": rb / (t["node_ms"] * 1e-3) / 1e9, "frac": rb / (t["node_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "edge_flux_evals_per_s": t["flux_evals"] / (t["node_ms"] * 1e-3),
            "traffic": measured_traffic(kname, "pmc_traffic_batch256.json") if len(meshes) == 256 else None}
    return out


def sweep_kernel_name(prec, tuning):
    """The sweep kernel a tuned engine of this precision launches (as rocprofv3 names it)."""
    import fsutil as U
    fsm = U.fsm
    tf = lambda v: "true" if v else "false"
    pre = tuning["gs32_pre"]
    if prec == fsm.FS_PREC_F64:
        if tuning.get("gs_lanes", 4) != 4:
            return "fsk::k_gs_row"
        return "fsk::k_gs_color<%d, %s>" % (pre, tf(tuning["gs32_dpp"] and pre > 0))
    pre = max(pre, 4)
    if prec == fsm.FS_PREC_MIXED16:
        return "fsk::k_gs_color16<%d, %s>" % (pre, tf(tuning["gs32_dpp"]))
    return "fsk::k_gs_color32<%d, %s, %s>" % (pre, tf(tuning["gs32_scaled"]), tf(tuning["gs32_dpp"]))


def batch_steady(a, fsm, cases, meshes, params, infos, prec):
    """Every solve to a 1e-10 drop under per-instance convergence masks
    (hipGraph steps); returns the outcome and the converged densities."""
    import numpy as np
    b = fsm.Batch(meshes, params, prec)
    b.use_graph(True)
    b.init_freestream()
    apply_controls(b)
    t0 = time.perf_counter()
    steps, drop, status = b.solve_steady(a.batch_max_steps, 1e-10)
    wall = time.perf_counter() - t0
    conv = status == fsm.INST_CONVERGED
    rec = np.array([b.recoveries(i)[0] for i in range(len(meshes))])
    fbits = np.array([b.recoveries(i)[1] for i in range(len(meshes))])
    fallback = fbits != 0
    rho = {i: b.get(i, "Q")[0:4 * infos[i]["nn"]:4] for i in range(len(meshes)) if conv[i]}
    fell = {i for i in range(len(meshes)) if fallback[i]}
    b.close()
    res = {"wall_s": wall, "converged": int(conv.sum()), "failed": int((status == fsm.INST_FAILED).sum()),
           "max_steps": int((status == fsm.INST_MAX_STEPS).sum()),
           "steps_median": float(np.median(steps[conv])) if conv.any() else None,
           "instance_steps_total": int(steps.sum()), "solves_per_s": float(conv.sum()) / wall,
           "converged_fraction": float(conv.mean()),
           "recovered_solves": int((rec > 0).sum()), "restores": int(rec.sum()),
           "converged_after_restore": int((conv & (rec > 0)).sum()),
           "limiter_fallbacks": int(((fbits & fsm.FALLBACK_LIMITER) != 0).sum()),
           "first_order_fallbacks": int(((fbits & fsm.FALLBACK_FIRST_ORDER) != 0).sum()),
           "stall_fallbacks": int(((fbits & fsm.FALLBACK_STALL) != 0).sum()),
           "converged_with_fallback": int((conv & fallback).sum()),
           "outcomes": outcome_groups(fsm, cases, status),
           "unconverged_mach_min": min([cases[i]["mach"] for i in range(len(cases)) if not conv[i]], default=None)}
    return res, rho, steps, fell


def plain_sequential(fsm, meshes, params, prec, warm, steps):
    """Each solve alone on the single-mesh solver (hipGraph steps): the mean
    wall time of its steps 2..warm+steps (or up to the step at which it stops
    on a non-physical state, as the batch stops it)."""
    ms, evals = [], 0
    for m, p in zip(meshes, params):
        s = fsm.Solver(m, p, prec)
        s.use_graph(True)
        s.init_freestream()
        times = []
        for ts in range(1, warm + steps + 1):
            t0 = time.perf_counter()
            try:
                s.solve(ts)
            except fsm.FlowSolverError:
                break
            if ts > 1:
                times.append(time.perf_counter() - t0)
        s.close()
        if not times:
            ms.append(0.0)
            continue
        ms.append(1e3 * sum(times) / len(times))
        evals += m.info()["nsub"] + m.info()["nbe"]
    live = sum(1 for v in ms if v > 0)
    return {"value": evals / (sum(ms) * 1e-3), "unit": "edge-flux evals/s", "ms_per_step_sum": sum(ms),
            "instances": live, "ms_per_step_per_instance": ms}


def bench_serving(a, precision):
    """Poisson arrivals at a fixed offered rate; each request's mesh is built
    before the clock starts and handed to the server at its arrival time (the
    server prepares it on its workers).  Latency = done - submit.  Requests
    still unfinished `serve_drain` seconds after the last arrival are
    cancelled and count as never converged."""
    import threading
    import numpy as np
    import workloads as W
    fsm = W.fsm
    rng = np.random.Generator(np.random.PCG64(77))
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / a.serve_rate)
        if t > a.serve_seconds:
            break
        arrivals.append(t)
    cases = W.serving_cases(len(arrivals))
    params = workload_params(W, cases)
    meshes = [W.make_mesh(c) for c in cases]
    progress("serving: %d request meshes built" % len(meshes))
    srv = fsm.Server(precision=precision, max_steps=a.serve_max_steps, tol=1e-10, prep_threads=16,
                     max_recoveries=RECOVERY[0], checkpoint_every=RECOVERY[1], limiter_fallback=RECOVERY[2],
                     first_order_fallback=RECOVERY[3], stall_steps_per_sqrt_node=STALL_STEPS_PER_SQRT_NODE,
                     max_steps_per_sqrt_node=SERVE_STEPS_PER_SQRT_NODE)
    res = []

    def collect():
        while len(res) < len(cases):
            info = srv.wait(600.0)
            if info is None:
                break
            res.append(info)
            srv.release(info["id"])

    th = threading.Thread(target=collect)
    th.start()
    w0 = time.perf_counter()
    name = {fsm.FS_PREC_MIXED: "mixed", fsm.FS_PREC_MIXED16: "mixed16"}.get(precision, "fp64")
    last = 0.0
    rid_case = {}
    for k, (arr, m, p) in enumerate(zip(arrivals, meshes, params)):
        d = arr - (time.perf_counter() - w0)
        if d > 0:
            time.sleep(d)
        rid_case[srv.submit(m, p)] = k
        now = time.perf_counter() - w0
        if now - last >= 60.0:
            last = now
            progress("serving %s: %d of %d submitted, %d finished at %.0f s" % (name, k + 1, len(cases), len(res), now))
    t_end = time.perf_counter() - w0
    cancelled = 0
    while th.is_alive():
        left = a.serve_drain - (time.perf_counter() - w0 - t_end)
        if left <= 0:
            # the drain deadline: what is still queued or running is cancelled
            done_ids = {r["id"] for r in list(res)}
            for rid in rid_case:
                if rid not in done_ids:
                    try:
                        srv.cancel(rid)
                        cancelled += 1
                    except fsm.FlowSolverError:
                        pass   # finished (and released) meanwhile
            th.join()
            break
        th.join(min(60.0, left))
        if th.is_alive():
            progress("serving %s: draining, %d of %d finished at %.0f s" %
                     (name, len(res), len(cases), time.perf_counter() - w0))
    wall = time.perf_counter() - w0
    srv.close()
    st = np.array([r["status"] for r in res])
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    dev = np.array([r["t_done"] - r["t_admit"] for r in res])
    conv = st == fsm.INST_CONVERGED
    progress("serving %s done: %d requests in %.1f s" % (name, len(res), wall))
    q = lambda v, p: float(np.percentile(v, p)) if len(v) else None

    def censored(v, ok, p):
        """p-th percentile (nearest rank) of the time to converged over every
        request, an unconverged request counting as never converged: None
        when the rank falls on one of those."""
        if not len(v):
            return None
        w = np.sort(np.where(ok, v, np.inf))
        r = w[max(0, int(np.ceil(p / 100.0 * len(w))) - 1)]
        return float(r) if np.isfinite(r) else None
    # outcomes by mesh kind and reconstruction order, and the Mach numbers of
    # the requests that did not converge
    groups = outcome_groups(fsm, [cases[rid_case[r["id"]]] for r in res], [r["status"] for r in res])
    bad = [cases[rid_case[r["id"]]]["mach"] for r in res if r["status"] != fsm.INST_CONVERGED]
    return {"requests": len(res), "offered_rate_per_s": a.serve_rate, "stream_s": a.serve_seconds, "wall_s": wall,
            "drain_deadline_s": a.serve_drain,
            "cancelled_at_deadline": int((st == fsm.INST_CANCELLED).sum()),
            "outcomes": groups, "unconverged_mach_min": min(bad) if bad else None,
            "converged": int(conv.sum()), "max_steps": int((st == fsm.INST_MAX_STEPS).sum()),
            "failed": int((st == fsm.INST_FAILED).sum()),
            "converged_fraction": float(conv.mean()) if len(conv) else None,
            "converged_fraction_of_finished": float(conv.sum()) / max(1, int((st != fsm.INST_CANCELLED).sum())),
            "recovered_requests": int(sum(1 for r in res if r["recoveries"] > 0)),
            "restores": int(sum(r["recoveries"] for r in res)),
            "limiter_fallbacks": int(sum(1 for r in res if r["fallbacks"] & fsm.FALLBACK_LIMITER)),
            "first_order_fallbacks": int(sum(1 for r in res if r["fallbacks"] & fsm.FALLBACK_FIRST_ORDER)),
            "stall_fallbacks": int(sum(1 for r in res if r["fallbacks"] & fsm.FALLBACK_STALL)),
            "step_limit_max": int(max(r["max_steps"] for r in res)) if res else None,
            "p50_time_to_converged_s": censored(lat, conv, 50), "p99_time_to_converged_s": censored(lat, conv, 99),
            "p50_admit_to_converged_s": censored(dev, conv, 50), "p99_admit_to_converged_s": censored(dev, conv, 99),
            "latency_note": "percentiles over all requests, an unconverged one counted as never converging (null "
                            "when the rank falls on one); *_converged_only over the converged requests",
            "p50_converged_only_s": q(lat[conv], 50), "p99_converged_only_s": q(lat[conv], 99),
            "solves_per_s": float(conv.sum()) / wall,
            "nodes_median": float(np.median([r["nn"] for r in res])) if res else None}


def _config3_solves(a, fsm, mesh, p, info):
    """FP64 and mixed steady solves of one config-3 mesh to a 1e-10 drop
    (hipGraph steps); the mixed path's error against the FP64 path."""
    import numpy as np
    out, rho, forces = {}, {}, {}
    evals = info["nsub"] + info["nbe"]
    for prec, name in ((fsm.FS_PREC_F64, "fp64"), (fsm.FS_PREC_MIXED, "mixed"), (fsm.FS_PREC_MIXED16, "mixed16")):
        s = fsm.Solver(mesh, p, prec)
        s.use_graph(True)
        s.init_freestream()
        t0 = time.perf_counter()
        try:
            steps, drop = s.solve_steady(a.config3_max_steps, 1e-10)
        except fsm.FlowSolverError as e:
            # a non-physical state: reported, and the section goes on
            out[name] = {"failed": str(e), "wall_s": time.perf_counter() - t0}
            progress("config3 %s failed: %s" % (name, e))
            s.close()
            continue
        wall = time.perf_counter() - t0
        forces[name] = s.body_forces(steps)
        rho[name] = s.get("Q")[0:4 * info["nn"]:4]
        t = s.time_steps(steps + 1, 10)
        out[name] = {"steps": steps, "drop": drop, "converged": bool(drop <= 1e-10), "wall_s": wall,
                     "ms_per_step": t["step_ms"], "value": evals / (t["step_ms"] * 1e-3), "unit": "edge-flux evals/s",
                     "lift": float(forces[name][0]), "drag": float(forces[name][1])}
        s.close()
        progress("config3 %s: %d steps, drop %.2e, %.1f s" % (name, steps, drop, wall))
    if "fp64" not in rho or "mixed" not in rho:
        return out
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["time_to_solution_speedup"] = out["fp64"]["wall_s"] / out["mixed"]["wall_s"]
    out["density_l2_diff_mixed_vs_fp64"] = float(np.linalg.norm(rho["mixed"] - rho["fp64"]) /
                                                 np.linalg.norm(rho["fp64"]))
    rel = np.abs(forces["mixed"] - forces["fp64"]) / np.abs(forces["fp64"])
    out["lift_rel_diff_mixed_vs_fp64"] = float(rel[0])
    out["drag_rel_diff_mixed_vs_fp64"] = float(rel[1])
    if "mixed16" in rho:
        rel = np.abs(forces["mixed16"] - forces["fp64"]) / np.abs(forces["fp64"])
        out["mixed16_vs_fp64"] = {
            "speedup": out["mixed16"]["value"] / out["fp64"]["value"],
            "time_to_solution_speedup": out["fp64"]["wall_s"] / out["mixed16"]["wall_s"],
            "density_l2_diff": float(np.linalg.norm(rho["mixed16"] - rho["fp64"]) / np.linalg.norm(rho["fp64"])),
            "lift_rel_diff": float(rel[0]), "drag_rel_diff": float(rel[1])}
    return out


def bench_config3(a, fsm):
    """The precision configuration: an O-mesh of 513 wall x 1171 radial nodes
    around the 12%-thick profile (first cell 1e-3 chord, 30 chords, +-0.1
    jitter, about 1.2 M triangles), M=0.8 at 1.25 degrees, k=2 with the Barth
    limiter, with a polyline wall.  FP64 and mixed steady solves to a 1e-10
    drop; the mixed path's error against the FP64 path for the converged
    density and the lift-like wall pressure integrals (Compute_Body_Forces).
    Then the same on the analytic NACA 0012 surfaces (codes 17/19, 512 wall
    nodes: the upper/lower split needs a leading-edge node), and the
    geometry's effect on lift and drag."""
    p = fsm.Params.make(nsteps=0, iter=1, jufreq=1, gamma=1.4, alpha=1.25, mach=0.8, method=1, order=3, foits=10,
                        soits=10, limit=1, reorder=1, **CONFIG3_CONTROLS)
    t0 = time.perf_counter()
    mesh = fsm.Mesh.airfoil(513, 1171, 1e-3, 30.0, 0.1, 3)
    mesh.prepare(p)
    info = mesh.info()
    out = {"workload": "O-mesh 513x1171 around the 12%-thick profile, M=0.8, alpha=1.25, k=2, Barth",
           "nodes": info["nn"], "triangles": info["ntri"], "subedges": info["nsub"], "bedges": info["nbe"],
           "setup_s": time.perf_counter() - t0, "controls": dict(CONFIG3_CONTROLS, tol=1e-10)}
    if a.config3_steps > 0:
        # short run: the first k=2-phase steps, launched one kernel at a time
        # (no hipGraph), with kernel times from events
        warm = p.foits + p.soits
        for prec, name in ((fsm.FS_PREC_F64, "fp64"), (fsm.FS_PREC_MIXED, "mixed"), (fsm.FS_PREC_MIXED16, "mixed16")):
            s = fsm.Solver(mesh, p, prec)
            s.init_freestream()
            for ts in range(1, warm + 1):
                s.solve(ts)
            t = s.time_steps(warm + 1, a.config3_steps)
            s.close()
            out[name] = {"ms_per_step": t["step_ms"], "value": t["flux_evals"] / (t["step_ms"] * 1e-3),
                         "unit": "edge-flux evals/s", "kernel_ms": {k: t[k] for k in ("edge_ms", "node_ms", "gs_ms",
                                                                                       "lu_ms", "other_ms")}}
        out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
        return out
    out.update(_config3_solves(a, fsm, mesh, p, info))
    del mesh
    if a.no_config3_curved:
        return out
    t0 = time.perf_counter()
    mesh = fsm.Mesh.airfoil_split(512, 1171, upper_bc=17, lower_bc=19, h0=1e-3, rmax=30.0, jitter=0.1, seed=3)
    mesh.prepare(p)
    info = mesh.info()
    cv = {"workload": "the same O-mesh with 512 wall nodes on the analytic NACA 0012 surfaces (codes 17/19)",
          "nodes": info["nn"], "triangles": info["ntri"], "setup_s": time.perf_counter() - t0}
    cv.update(_config3_solves(a, fsm, mesh, p, info))
    if "lift" in cv.get("fp64", {}) and "lift" in out.get("fp64", {}):
        cv["lift_curved_vs_polyline"] = cv["fp64"]["lift"] / out["fp64"]["lift"]
        cv["drag_curved_vs_polyline"] = cv["fp64"]["drag"] / out["fp64"]["drag"]
    out["curved"] = cv
    return out


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


def cpu_reference_calibration(U, fsm, nx, ny, steps=2):
    """The reference build itself (oracle/_ref/ref_driver, compiled from the
    reference's sources) on the same bump channel, 1 core, against the port
    on the same steps.  The reference build has no LAPACK, so it cannot enter
    the k=2 phase: both run `steps` steps of the limited linear phase
    (order 1 from the first step), which has the k=2 phase's FD Jacobian and
    20 sequential sweeps plus the gradient and the limiter.  None when the
    reference build is absent."""
    import subprocess
    import tempfile
    if not os.path.exists(U.REF_DRIVER):
        return None
    p = fsm.Params.make(nsteps=steps, iter=1, jufreq=1, CFL=50.0, ramp=20, gamma=1.4, alpha=0.0, mach=0.675,
                        subits=20, method=1, order=1, foits=0, soits=0, limit=1, reorder=1)
    mesh = fsm.Mesh.bump(nx, ny, 0.2, 2)
    with tempfile.TemporaryDirectory() as tmp:
        U.write_case(tmp, mesh, p)
        r = subprocess.run([U.REF_DRIVER, tmp, os.path.join(tmp, "dump.bin"), str(steps), "0", "0"],
                           capture_output=True, text=True, cwd=tmp, timeout=600)
        if r.returncode != 0:
            return {"error": r.stderr[-300:]}
        t_ref = float(U.read_dump(os.path.join(tmp, "dump.bin"))["times"][1])
    mesh.prepare(p)
    info = mesh.info()
    o = U.Oracle(mesh, p)
    o.init_freestream()
    t0 = time.perf_counter()
    for ts in range(1, steps + 1):
        o.solve(ts, multicolor=0)
    t_port = time.perf_counter() - t0
    evals = (info["nsub"] + info["nbe"]) * steps
    return {"reference_value": evals / t_ref, "port_value": evals / t_port, "unit": "edge-flux evals/s",
            "cores": 1, "reference_over_port": t_port / t_ref,
            "sample": "%d limited linear-phase steps of the %dx%d bump channel: reference %.1f s, port %.1f s"
                      % (steps, nx, ny, t_ref, t_port)}


def _cpu_batch_worker(case, first, steps, barrier, queue, core):
    """One batch solve on the oracle in its own process (spawned: no GPU),
    pinned to one core: mesh preparation untimed, the state set to the
    freestream perturbed by 5% (seeded), then `steps` pseudo-time steps of
    the final phase (from step `first`) timed after a barrier shared by all
    workers."""
    os.sched_setaffinity(0, {core})
    import fsutil as U
    import workloads as W
    p = workload_params(W, [case])[0]
    m = W.make_mesh(case)
    m.prepare(p)
    info = m.info()
    o = U.Oracle(m, p)
    o["Q"][:] = U.perturbed_state(info["nn"], info["ng"], p, seed=case["seed"])
    barrier.wait()
    t0 = time.time()
    for ts in range(first, first + steps):
        o.solve(ts, multicolor=0)
    queue.put((t0, time.time(), (info["nsub"] + info["nbe"]) * steps, info["ntri"]))


def cpu_baseline_all_cores(first, steps=2):
    """The oracle on every host core this process may use (at most 16): one
    independent solve of the headline batch per core (each process pinned to
    its core), its final-phase steps timed together after a barrier.
    Throughput = the sum of the per-core rates (a CPU deployment keeps every
    core busy with a queue of solves); the span from the first start to the
    last finish is reported beside it."""
    import multiprocessing as mp
    import workloads as W
    cores = sorted(os.sched_getaffinity(0))[:16]
    ncores = len(cores)
    cases = W.batch_cases(ncores)
    ctx = mp.get_context("spawn")
    barrier, queue = ctx.Barrier(ncores), ctx.Queue()
    procs = [ctx.Process(target=_cpu_batch_worker, args=(c, first, steps, barrier, queue, core))
             for c, core in zip(cases, cores)]
    for pr in procs:
        pr.start()
    res = []
    while len(res) < len(procs):
        try:
            res.append(queue.get(timeout=5))
        except Exception:
            if any(pr.exitcode not in (None, 0) for pr in procs):
                for pr in procs:
                    if pr.is_alive():
                        pr.terminate()
                raise RuntimeError("a CPU baseline worker failed")
    for pr in procs:
        pr.join()
    span = max(r[1] for r in res) - min(r[0] for r in res)
    rate = sum(r[2] / (r[1] - r[0]) for r in res)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": rate, "unit": "edge-flux evals/s", "cores": ncores, "kind": "port", "cpu": model,
            "value_over_span": sum(r[2] for r in res) / span,
            "sample": "%d final-phase pseudo-time steps (from step %d, the timed phase of the GPU run) of each of the "
                      "first %d solves of the same 256-solve workload (%d triangles in all), one process pinned to "
                      "each core, %.1f s" % (steps, first, ncores, sum(r[3] for r in res), span)}


def bench_config2(a, fsm):
    """The 1025x257 bump channel at M=0.675, k=2, Barth: FP64 and mixed steps
    of the k=2 phase, with kernel times from events."""
    p = fsm.Params.make(nsteps=0, iter=1, jufreq=1, CFL=50.0, ramp=20, gamma=1.4, alpha=0.0, mach=0.675, subits=20,
                        method=1, order=3, foits=2, soits=2, limit=1, reorder=1)
    t0 = time.perf_counter()
    mesh = fsm.Mesh.bump(a.nx, a.ny, 0.2, 2)
    mesh.prepare(p)
    t_setup = time.perf_counter() - t0
    info = mesh.info()
    evals = info["nsub"] + info["nbe"]
    warm = max(a.config2_warmup, p.foits + p.soits)
    out = {"workload": "bump channel %dx%d, M=0.675, k=2 (order 3), Barth limiter" % (a.nx, a.ny),
           "nodes": info["nn"], "subedges": info["nsub"], "bedges": info["nbe"], "lhs_blocks": info["nnz"],
           "subits": p.subits, "steps": a.config2_steps, "warmup": warm, "setup_s": t_setup,
           "jacobian_flux_evals_per_step_fp64": 8 * info["nsub"]}
    for prec, name in ((fsm.FS_PREC_F64, "fp64"), (fsm.FS_PREC_MIXED, "mixed"), (fsm.FS_PREC_MIXED16, "mixed16")):
        s = fsm.Solver(mesh, p, prec)
        s.init_freestream()
        s.solve(1)
        s.autotune()   # launch shapes on this mesh (results unchanged)
        for ts in range(2, warm + 1):
            s.solve(ts)
        t = s.time_steps(warm + 1, a.config2_steps)
        tuning = s.tuning()
        scaled = tuning["gs32_scaled"]
        s.close()
        word = 8 if prec == fsm.FS_PREC_F64 else 4
        block_word = 2 if prec == fsm.FS_PREC_MIXED16 else word
        gs_bytes, gs_launches = algorithmic_bytes_gauss_seidel(mesh, p.subits, word, scaled, block_word)
        r = {"value": evals / (t["step_ms"] * 1e-3), "unit": "edge-flux evals/s", "ms_per_step": t["step_ms"],
             "gauss_seidel": {"kernel": sweep_kernel_name(prec, tuning), "launches_per_step": gs_launches,
                              "avg_ms": t["gs_ms"] / gs_launches,
                              "algorithmic_bytes_per_launch": gs_bytes / gs_launches,
                              "achieved_GBs": gs_bytes / (t["gs_ms"] * 1e-3) / 1e9}}
        if prec == fsm.FS_PREC_F64:
            ab = algorithmic_bytes_edge_kernel(info, recon=0)
            r["kernel_ms"] = {"edge_flux_fd_jacobian": t["edge_ms"], "node_assembly": t["node_ms"],
                              "gauss_seidel": t["gs_ms"], "diag_lu": t["lu_ms"], "other": t["other_ms"]}
            r["edge_kernel"] = {"kernel": "k_edge_flux (Roe flux + FD Jacobian blocks, FP64 VALU)",
                                "avg_ms": t["edge_ms"], "algorithmic_bytes": ab,
                                "achieved_GBs": ab / (t["edge_ms"] * 1e-3) / 1e9}
            r["gauss_seidel"]["traffic"] = measured_traffic(r["gauss_seidel"]["kernel"])
        else:
            fused = 1 if prec == fsm.FS_PREC_MIXED16 or (scaled and tuning["node_fused"]) else 0
            kname = "fsk::k_resid_jac32<%d>" % (2 if prec == fsm.FS_PREC_MIXED16 else fused)
            rb = algorithmic_bytes_resid_jac32(info, 0, fused, block_word)
            r["kernel_ms"] = {"boundary": t["edge_ms"], "residual_jacobian_node_pass": t["node_ms"],
                              "gauss_seidel": t["gs_ms"], "diag_inverse": t["lu_ms"], "other": t["other_ms"]}
            r["residual_kernel"] = {"kernel": kname, "avg_ms": t["node_ms"], "algorithmic_bytes": rb,
                                    "achieved_GBs": rb / (t["node_ms"] * 1e-3) / 1e9,
                                    "frac": rb / (t["node_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                    "traffic": measured_traffic(kname)}
            r["gauss_seidel"]["traffic"] = measured_traffic(r["gauss_seidel"]["kernel"])
        out[name] = r
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["speedup_mixed16_vs_fp64"] = out["mixed16"]["value"] / out["fp64"]["value"]
    return out, mesh, p


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fsutil as U

    fsm = U.fsm
    sections = set(a.sections.split(",")) - {""}
    if a.skip_headline:
        line = {"metric": "edge-flux evals/s + steady-solve p50/p99 latency, mixed vs FP64, 1 GPU", "value": None,
                "note": "headline not run (--skip-headline)"}
        if "config2" in sections:
            line["config2"], _, _ = bench_config2(a, fsm)
        if "config3" in sections:
            line["config3"] = bench_config3(a, fsm)
        print(json.dumps(line), flush=True)
        return
    cases, params, meshes, infos, desc = build_batch(a)
    progress("batch built: %d instances, %d nodes (%.1f s)" % (len(meshes), desc["nodes"], desc["setup_s"]))
    mixed = batch_throughput(a, fsm, meshes, params, infos, fsm.FS_PREC_MIXED)
    gs = mixed["gauss_seidel"]
    progress("headline (mixed batch): %.3f ms/step" % mixed["ms_per_step"])
    line = {
        "metric": "edge-flux evals/s + steady-solve p50/p99 latency, mixed vs FP64, 1 GPU",
        "value": mixed["value"], "unit": "edge-flux evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": mixed["warmup"], "ms_per_step": mixed["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded meshes and flow conditions, freestream start)",
        "config": dict({"workload": "256 independent steady solves in one engine (bump channels and airfoil "
                                    "O-meshes, 5k-200k triangles log-uniform, M~U[0.3,0.8], alpha~U[-2,4] deg, "
                                    "k in {1,2}), mixed precision (FP64 residual, norms and update; FP32 analytic "
                                    "Jacobian and Gauss-Seidel sweeps), one launch per kernel per pseudo-time step, "
                                    "final-phase steps",
                        "precision": "mixed"}, **desc),
        "roofline": {"bound": "hbm", "kernel": "%s (multicolor block Gauss-Seidel sweep, FP32 blocks; "
                                               "%.0f%% of the step)" % (gs["kernel"], 100 * mixed["kernel_ms"][
                                                   "gauss_seidel"] / mixed["ms_per_step"]),
                     "achieved": gs["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gs["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(gs["kernel"], "pmc_traffic_batch256.json")
                     if a.batch == 256 else None,
                     "algorithmic_bytes": gs["algorithmic_bytes_per_launch"],
                     "launches_per_step": gs["launches_per_step"], "avg_ms": gs["avg_ms"]},
        "residual_kernel": mixed["residual_kernel"],
        "mixed": mixed,
    }
    rd, cp = fsm.hbm_probe(4 << 30, 5)
    line["achievable"] = {"read_GBs": rd, "copy_GBs": cp, "gs_frac_of_read": gs["achieved_GBs"] / rd,
                          "probe": "fs_hbm_probe: streaming float4 read and copy of 4 GiB, 5 passes each"}
    if "fp64" in sections:
        f = batch_throughput(a, fsm, meshes, params, infos, fsm.FS_PREC_F64)
        line["fp64"] = f
        line["speedup_mixed_vs_fp64"] = mixed["value"] / f["value"]
        progress("fp64 batch: %.3f ms/step" % f["ms_per_step"])
    if "mixed16" in sections:
        # the same path with the sweep operands stored in FP16 (FP32 sweep
        # arithmetic): an option beside the headline, not the headline
        h = batch_throughput(a, fsm, meshes, params, infos, fsm.FS_PREC_MIXED16)
        h["roofline"] = {"kernel": h["gauss_seidel"]["kernel"], "achieved": h["gauss_seidel"]["achieved_GBs"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": h["gauss_seidel"]["achieved_GBs"] / HBM_PEAK_GBS,
                         "traffic": measured_traffic(h["gauss_seidel"]["kernel"], "pmc_traffic_batch256.json")
                         if a.batch == 256 else None}
        line["mixed16"] = h
        line["speedup_mixed16_vs_mixed"] = h["value"] / mixed["value"]
        if "fp64" in line:
            line["speedup_mixed16_vs_fp64"] = h["value"] / line["fp64"]["value"]
        progress("mixed16 batch: %.3f ms/step" % h["ms_per_step"])
    rho, steps = {}, {}
    if "steady" in sections and a.batch_max_steps > 0:
        st, fell, fell16 = {}, set(), set()
        runs = [(fsm.FS_PREC_MIXED, "mixed"), (fsm.FS_PREC_F64, "fp64")]
        if "mixed16" in sections:
            runs.insert(1, (fsm.FS_PREC_MIXED16, "mixed16"))
        for prec, name in runs:
            st[name], rho[name], steps[name], f = batch_steady(a, fsm, cases, meshes, params, infos, prec)
            if name == "mixed16":
                fell16 = f
            else:
                fell |= f
            progress("steady %s: %d converged in %.1f s" % (name, st[name]["converged"], st[name]["wall_s"]))
        if "mixed16" in rho:
            # FP16 sweep operands: the same comparisons against the FP64 path
            b16 = sorted(set(rho["fp64"]) & set(rho["mixed16"]))
            d16 = [float(np.linalg.norm(rho["mixed16"][i] - rho["fp64"][i]) / np.linalg.norm(rho["fp64"][i]))
                   for i in b16 if i not in fell | fell16]
            st["mixed16_vs_fp64"] = {"converged_both": len(b16), "compared": len(d16),
                                     "max_density_l2_diff": max(d16) if d16 else None,
                                     "steps_total_mixed16": int(sum(steps["mixed16"][i] for i in b16)),
                                     "steps_total_fp64": int(sum(steps["fp64"][i] for i in b16)),
                                     "time_to_solution_speedup": st["fp64"]["wall_s"] / st["mixed16"]["wall_s"]}
        both = sorted(set(rho["fp64"]) & set(rho["mixed"]))
        diff = {i: float(np.linalg.norm(rho["mixed"][i] - rho["fp64"][i]) / np.linalg.norm(rho["fp64"][i]))
                for i in both}
        plain_ids = [i for i in both if i not in fell]
        fb_ids = [i for i in both if i in fell]
        st["converged_both"] = len(both)
        # solves on their own discretization: the mixed path's error against FP64
        st["max_density_l2_diff_mixed_vs_fp64"] = max(diff[i] for i in plain_ids) if plain_ids else None
        st["compared"] = len(plain_ids)
        # solves the limiter fallback touched in either precision: frozen
        # limiter values differ between the two iterate paths
        st["fallback_solves_compared"] = len(fb_ids)
        st["max_density_l2_diff_fallback"] = max(diff[i] for i in fb_ids) if fb_ids else None
        st["time_to_solution_speedup"] = st["fp64"]["wall_s"] / st["mixed"]["wall_s"]
        line["steady"] = st
    if "plain" in sections:
        # the plain FP64 GPU path: the same solves one at a time (no
        # batching), each timed alone over steps of its final phase
        warm = max(p.foits + p.soits for p in params)
        plain = plain_sequential(fsm, meshes, params, fsm.FS_PREC_F64, warm, 4)
        line["speedup_mixed_batched_vs_plain_fp64"] = mixed["value"] / plain["value"]
        if "fp64" in steps:
            # an instance's FP64 trajectory is the same alone and batched, so
            # the plain path needs the same steps
            t_plain = sum(n * ms for n, ms in zip(steps["fp64"], plain["ms_per_step_per_instance"])) * 1e-3
            plain["time_to_solution_s_estimated"] = t_plain
            line["time_to_solution_speedup_vs_plain_fp64"] = t_plain / line["steady"]["mixed"]["wall_s"]
        plain.pop("ms_per_step_per_instance", None)
        line["plain_fp64_sequential"] = plain
        progress("plain fp64 done")
    del meshes
    if "config2" in sections:
        line["config2"], mesh2, p2 = bench_config2(a, fsm)
        progress("config2 done")
    if "config3" in sections:
        line["config3"] = bench_config3(a, fsm)
    if "serving" in sections and a.serve_seconds > 0:
        precs = {"mixed": fsm.FS_PREC_MIXED, "fp64": fsm.FS_PREC_F64, "mixed16": fsm.FS_PREC_MIXED16}
        sv = {k: bench_serving(a, precs[k]) for k in a.serve_precisions.split(",") if k}
        sv["controls"] = dict(workload_controls(), max_steps=a.serve_max_steps,
                              max_steps_per_sqrt_node=SERVE_STEPS_PER_SQRT_NODE)
        line["serving"] = sv
    if "cpu" in sections and not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_all_cores(mixed["warmup"] + 1)
        if "config2" in sections:
            line["cpu_baseline_config2"] = cpu_baseline(U, fsm, p2, a.nx, a.ny, a.cpu_steps)
        line["cpu_reference_calibration"] = cpu_reference_calibration(U, fsm, a.nx, a.ny)
        progress("cpu baselines done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
