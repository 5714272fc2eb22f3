"""Batched solves on the GPU: instances of different meshes, Mach numbers,
angles, reconstruction orders and limiter settings packed into one launch per
kernel.  An instance's arithmetic does not depend on its neighbours, so every
result must equal the same solve run alone bit for bit (a stronger bar than
the 1e-10 tolerance the serving design asks for).  Also: graph replay equals
eager launches, and a masked-off instance is left untouched."""
import numpy as np
import pytest

import fsutil as U

fsm = U.fsm
pytestmark = pytest.mark.gpu


def _params(**kw):
    base = dict(nsteps=0, iter=1, jufreq=1, CFL=100.0, ramp=30, gamma=1.4, alpha=0.0, mach=0.5, subits=12,
                method=1, order=1, foits=5, soits=0, limit=0, reorder=1)
    base.update(kw)
    return fsm.Params.make(**base)


def _cases():
    """(mesh factory, params): uneven sizes, both orders, straight and curved walls."""
    return [
        (lambda: fsm.Mesh.bump(33, 17, 0.2, 1), _params(mach=0.5, order=1)),
        (lambda: fsm.Mesh.bump(41, 13, 0.2, 2), _params(mach=0.6, alpha=1.0, order=3, soits=4, limit=1)),
        (lambda: fsm.Mesh.airfoil(48, 10, 2e-2, 8.0, 0.05, 3), _params(mach=0.4, alpha=2.0, order=1)),
        # 3,201 nodes: four norm chunks, at a non-zero position in the pack
        (lambda: fsm.Mesh.bump(97, 33, 0.2, 1), _params(mach=0.5, alpha=0.5, order=3, soits=3, limit=1)),
        (lambda: fsm.Mesh.bump(17, 9, 0.1, 4), _params(mach=0.3, alpha=-1.0, order=3, soits=2, limit=0)),
        (lambda: fsm.Mesh.bump(25, 11, 0.2, 5), _params(mach=0.55, order=1, limit=1, foits=3)),
        # 3 colors where the others have 4: its backward launches of color 2
        # must be skipped for it to see its own sweep order
        (lambda: fsm.Mesh.bump(3, 3, 0.0, 1), _params(mach=0.4, order=1)),
        # analytic curved walls: NACA 0012 surfaces, cylinder in a circular farfield
        (lambda: fsm.Mesh.airfoil_split(40, 9, upper_bc=17, lower_bc=19, h0=2e-2, rmax=8.0, jitter=0.1, seed=6),
         _params(mach=0.5, alpha=1.5, order=3, soits=2, limit=1)),
        (lambda: fsm.Mesh.annulus(32, 9, r_in=1.0, r_out=300.0, bc_in=11, bc_out=22, jitter=0.1, seed=7),
         _params(mach=0.3, alpha=0.5, order=1)),
        # triangles and quadrilaterals
        (lambda: U.bump_mixed(29, 11, 8), _params(mach=0.5, order=3, soits=2, limit=1)),
    ]


def _build(precision):
    cases = _cases()
    meshes = []
    for make, p in cases:
        m = make()
        m.prepare(p)
        meshes.append(m)
    return meshes, [p for _, p in cases]


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED, fsm.FS_PREC_MIXED16])
def test_batch_equals_alone(precision):
    meshes, params = _build(precision)
    print("case shapes", [(m.info()["nn"], m.info()["ncolor"]) for m in meshes], flush=True)
    assert len({m.info()["ncolor"] for m in meshes}) > 1
    b = fsm.Batch(meshes, params, precision)
    b.init_freestream()
    steps, drop, status = b.solve_steady(120, 1e-6)
    print("batch steps", steps, "drop", drop, "status", status, flush=True)
    forces = b.body_forces()
    for i, (m, p) in enumerate(zip(meshes, params)):
        s = fsm.Solver(m, p, precision)
        s.init_freestream()
        st, dr = s.solve_steady(120, 1e-6)
        # Compute_Body_Forces of the final state, alone and in the batch
        np.testing.assert_array_equal(s.body_forces(st), forces[i])
        same = np.array_equal(s.get("Q"), b.get(i, "Q"))
        print("instance %d alone steps %d drop %.3e  batch steps %d drop %.3e  bit-identical %s" %
              (i, st, dr, steps[i], drop[i], same), flush=True)
        assert st == steps[i]
        assert dr == drop[i]
        assert same
        assert np.array_equal(s.get("R"), b.get(i, "R"))
        s.close()
    # the convergence mask removed instances at different steps
    assert status.tolist().count(fsm.INST_CONVERGED) >= 1
    assert len(set(steps.tolist())) > 1


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED, fsm.FS_PREC_MIXED16])
def test_deferred_k2_fields_equal_every_step(precision):
    """The k=2 phase's derivatives, point values and limiter values deferred
    to the outputs that read them (the default) against evaluated every step
    as the reference does: the residual norms and states of every step, the
    fields read two steps before the end (the deferred run then goes on
    deferring), and the body forces and fields at the end, bit for bit.  The
    k=2 instances reach their k=2 phase by step 10 (foits 5, soits at most
    4)."""
    meshes, params = _build(precision)
    n = len(meshes)
    nsteps = 16
    runs = {}
    for defer in (False, True):
        b = fsm.Batch(meshes, params, precision)
        b.init_freestream()
        b.use_graph(True)
        b.defer_k2(defer)
        hist, mid = [], None
        for ts in range(1, nsteps + 1):
            r2, dq = b.step()
            hist.append((r2, dq, [b.get(i, "Q") for i in range(n)]))
            if ts == nsteps - 2:
                mid = [[b.get(i, f) for f in ("phi", "hess", "pv")] for i in range(n)]
        forces = b.body_forces()
        end = [[b.get(i, f) for f in ("phi", "hess", "pv")] for i in range(n)]
        runs[defer] = (hist, mid, forces, end)
        b.close()
    (h0, m0, f0, e0), (h1, m1, f1, e1) = runs[False], runs[True]
    for ts in range(nsteps):
        assert np.array_equal(h0[ts][0], h1[ts][0]) and np.array_equal(h0[ts][1], h1[ts][1]), "norms, step %d" % (ts + 1)
        for i in range(n):
            assert np.array_equal(h0[ts][2][i], h1[ts][2][i]), "state of instance %d, step %d" % (i, ts + 1)
    for i in range(n):
        for k, f in enumerate(("phi", "hess", "pv")):
            assert np.array_equal(m0[i][k], m1[i][k]), "%s of instance %d, step %d" % (f, i, nsteps - 2)
            assert np.array_equal(e0[i][k], e1[i][k]), "%s of instance %d, end" % (f, i)
    assert np.array_equal(f0, f1)
    # the k=2 instances did run their k=2 phase, with limiter values below 1
    k2 = [i for i, p in enumerate(params) if p.order == 3 and p.limit == 1]
    assert k2 and all(np.min(e1[i][0]) < 1.0 for i in k2)


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED, fsm.FS_PREC_MIXED16])
def test_graph_replay_equals_eager(precision):
    p = _params(order=3, soits=2, limit=1, foits=2)
    m = fsm.Mesh.bump(33, 17, 0.2, 7)
    m.prepare(p)
    eager = fsm.Solver(m, p, precision)
    graph = fsm.Solver(m, p, precision)
    graph.use_graph(True)
    eager.init_freestream()
    graph.init_freestream()
    # phases change at ts 3 and 5: three graphs are captured and each is
    # replayed several times, with eager work and synchronisation in between
    for ts in range(1, 12):
        de, re = eager.solve(ts)
        dg, rg = graph.solve(ts)
        qe, qg = eager.get("Q"), graph.get("Q")
        print("ts %d eager |dQ| %.15e graph %.15e identical %s" % (ts, de, dg, np.array_equal(qe, qg)), flush=True)
        assert np.array_equal(qe, qg)
        assert de == dg and re == rg


def test_mask_freezes_instance():
    meshes, params = _build(fsm.FS_PREC_F64)
    meshes, params = meshes[:2], params[:2]
    b = fsm.Batch(meshes, params, fsm.FS_PREC_F64)
    b.use_graph(True)
    b.init_freestream()
    for _ in range(3):
        b.step()
    q1 = b.get(1, "Q")
    b.set_active(1, False)
    for _ in range(3):
        b.step()
    assert np.array_equal(b.get(1, "Q"), q1)
    assert b.status(1)[0] == fsm.INST_PAUSED
    # instance 0 kept stepping: equal to 6 steps alone
    s = fsm.Solver(meshes[0], params[0], fsm.FS_PREC_F64)
    s.init_freestream()
    for ts in range(1, 7):
        s.solve(ts)
    assert np.array_equal(s.get("Q"), b.get(0, "Q"))
    # resuming continues from the frozen state and step counter
    b.set_active(1, True)
    b.step()
    s1 = fsm.Solver(meshes[1], params[1], fsm.FS_PREC_F64)
    s1.init_freestream()
    for ts in range(1, 5):
        s1.solve(ts)
    assert np.array_equal(s1.get("Q"), b.get(1, "Q"))


def test_restart_continuation_bitexact(tmp_path):
    """Steps 1..5, restart file with the ghost states, then steps 6..10 from
    the file on a fresh mesh and solver: identical to ten uninterrupted steps.
    Without the ghost extension (the reference's own restart, which copies
    each real node into its ghosts, main.C:604-620) the run continues but the
    ghost memory of the characteristic boundary update is lost."""
    p = _params(order=1, limit=1, foits=2)
    m = fsm.Mesh.bump(33, 13, 0.2, 9)
    m.prepare(p)
    ref = fsm.Solver(m, p)
    ref.init_freestream()
    for ts in range(1, 6):
        ref.solve(ts)
    bc = str(tmp_path / "bc.dat")
    with open(bc, "w") as f:
        f.write("".join("%d\n" % v for v in m.array("bbc")))
    out = {}
    for ghosts in (True, False):
        path = str(tmp_path / ("restart_%d.mesh" % ghosts))
        m.write_restart(path, ref.get("Q"), 5, p.CFL, ghosts=ghosts)
        pr = fsm.Params.make(**dict(p.as_dict(), isrestart=1))
        m2 = fsm.Mesh.read(path, bc)
        m2.prepare(pr)
        s2 = fsm.Solver(m2, pr)
        s2.init_restart()
        for ts in range(6, 11):
            s2.solve(ts)
        out[ghosts] = s2.get("Q")
    for ts in range(6, 11):
        ref.solve(ts)
    q = ref.get("Q")
    print("continuation with ghost states identical: %s; without: max |dQ| %.2e" %
          (np.array_equal(out[True], q), np.max(np.abs(out[False] - q))))
    assert np.array_equal(out[True], q)
    assert np.all(np.isfinite(out[False]))


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED])
def test_autotune_keeps_results_and_caches(precision, tmp_path, monkeypatch):
    """Tuned launch shapes give the same states bit for bit; the choice is
    written to the cache and picked up by the next engine with the same key.
    The FP32 sweep's form (the one launch choice that changes FP32 rounding)
    is neither tuned nor cached: pinned either way, the tuned run keeps it."""
    cache = tmp_path / "tune_cache.txt"
    monkeypatch.setenv("FS_TUNE_CACHE", str(cache))
    meshes, params = _build(precision)

    def run(tune, scaled=None):
        b = fsm.Batch(meshes, params, precision)
        b.use_graph(True)
        if scaled is not None:
            b.set_tuning(gs32_scaled=scaled)
        b.init_freestream()
        b.step()
        info = b.autotune() if tune else b.tuning()
        for _ in range(8):
            b.step()
        q = [b.get(i, "Q") for i in range(len(meshes))]
        b.close()
        return q, info

    q0 = {sc: run(False, sc)[0] for sc in (0, 1)}
    assert not cache.exists() or cache.read_text() == ""
    q1, t1 = run(True, 0)
    print("tuned shapes", t1, flush=True)
    assert t1["gs32_scaled"] == 0
    assert cache.exists() and len(cache.read_text().splitlines()) == 1
    for a, b in zip(q0[0], q1):
        assert np.array_equal(a, b)
    if precision == fsm.FS_PREC_F64:
        # the FP64 path has no FP32 sweep: the knob is inert
        for a, b in zip(q0[0], q0[1]):
            assert np.array_equal(a, b)
    else:
        # the two forms round differently but solve the same system
        for a, b in zip(q0[0], q0[1]):
            assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(a))
    b = fsm.Batch(meshes, params, precision)
    t2 = b.tuning()
    b.close()
    assert t2["from_cache"]
    assert {k: t2[k] for k in t2 if k != "from_cache"} == {k: t1[k] for k in t2 if k != "from_cache"}


def test_nonphysical_instance_fails_alone():
    """A non-physical state (negative density at one node, set from the host)
    ends that solve with the reference's fatal condition as a status: the
    single-mesh solver returns FS_ERR_STATE naming the exit it replaces, and
    in a batch only that instance is marked failed while the others run on
    and still equal their solves alone."""
    meshes, params = _build(fsm.FS_PREC_F64)
    meshes, params = meshes[:3], params[:3]
    bad = 1
    info = meshes[bad].info()
    s = fsm.Solver(meshes[bad], params[bad])
    s.init_freestream()
    Q = s.get("Q")
    Q[4 * (info["nn"] // 2)] = -0.5
    s.set("Q", Q)
    with pytest.raises(fsm.FlowSolverError) as e:
        s.solve(1)
    print("single solver:", e.value, flush=True)
    assert e.value.code == -5
    assert "init_memory.C" in str(e.value) or "cvbc.C" in str(e.value)
    s.close()
    b = fsm.Batch(meshes, params)
    b.init_freestream()
    b.set(bad, "Q", Q)
    steps, drop, status = b.solve_steady(80, 1e-6)
    print("batch status", status, "steps", steps, flush=True)
    assert status[bad] == fsm.INST_FAILED and steps[bad] == 1
    for i in range(3):
        if i == bad:
            continue
        assert status[i] != fsm.INST_FAILED
        a = fsm.Solver(meshes[i], params[i])
        a.init_freestream()
        st, dr = a.solve_steady(80, 1e-6)
        assert st == steps[i]
        assert np.array_equal(a.get("Q"), b.get(i, "Q"))
        a.close()


def test_compaction_keeps_results():
    """Steady solves of 10 instances that stop at different steps: with the
    running instances re-packed into smaller engines as others stop, every
    instance's steps, drop, final state, residual and body forces equal those
    of the same batch without re-packing and of the solve alone."""
    cases = []
    for k in range(10):
        nx, ny = 17 + 4 * (k % 4), 9 + 2 * (k % 3)
        kw = dict(mach=0.3 + 0.04 * k, alpha=-1.0 + 0.5 * k, order=1 if k % 3 else 3, soits=0 if k % 3 else 3,
                  limit=0 if k % 3 else 1)
        cases.append(((nx, ny, 0.2, 40 + k), _params(**kw)))
    meshes = []
    for (nx, ny, amp, seed), p in cases:
        m = fsm.Mesh.bump(nx, ny, amp, seed)
        m.prepare(p)
        meshes.append(m)
    params = [p for _, p in cases]
    runs = {}
    for on in (True, False):
        b = fsm.Batch(meshes, params)
        b.set_compaction(on)
        b.init_freestream()
        steps, drop, status = b.solve_steady(200, 1e-8)
        runs[on] = (b, steps, drop, status, b.body_forces(), b.info()["device_bytes"])
    bc, steps, drop, status, forces, bytes_on = runs[True]
    bn, steps_n, drop_n, status_n, forces_n, bytes_off = runs[False]
    print("steps", steps, "status", status, "device bytes with / without re-packing", bytes_on, bytes_off, flush=True)
    assert len(set(steps.tolist())) > 3
    assert bytes_on > bytes_off, "no re-packing happened"
    np.testing.assert_array_equal(steps, steps_n)
    np.testing.assert_array_equal(drop, drop_n)
    np.testing.assert_array_equal(status, status_n)
    np.testing.assert_array_equal(forces, forces_n)
    for i in range(10):
        for f in ("Q", "R", "dQ"):
            assert np.array_equal(bc.get(i, f), bn.get(i, f)), (i, f)
    for i in (int(np.argmax(steps)), int(np.argmin(steps))):
        a = fsm.Solver(meshes[i], params[i])
        a.init_freestream()
        st, dr = a.solve_steady(200, 1e-8)
        assert st == steps[i] and dr == drop[i]
        assert np.array_equal(a.get("Q"), bc.get(i, "Q"))
        np.testing.assert_array_equal(a.body_forces(st), forces[i])
        a.close()


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED])
def test_derivative_lds_tiles_equal_gathers(precision):
    """The k=2 derivatives with each 32-node tile's neighbour states staged in
    LDS give the gather kernel's moments and point values bit for bit, over a
    pack whose tiles span instances (k=1 instances have empty stencils)."""
    meshes, params = _build(precision)
    res = []
    for lds in (0, 1):
        b = fsm.Batch(meshes, params, precision)
        b.set_tuning(deriv_lds=lds)
        assert b.tuning()["deriv_lds"] == lds
        b.init_freestream()
        for _ in range(12):
            b.step()
        res.append([(b.get(i, "hess"), b.get(i, "pv"), b.get(i, "Q")) for i in range(len(meshes))])
        b.close()
    k2 = [i for i, p in enumerate(params) if p.order == 3]
    assert any(np.any(res[1][i][0] != 0.0) for i in k2), "no instance reached the k=2 phase"
    for i, (a, c) in enumerate(zip(*res)):
        for x, y in zip(a, c):
            assert np.array_equal(x, y), "instance %d" % i


def test_fused_operands_equal_diagonal_pass():
    """With 40 sweeps the FP32 sweeps run on operands pre-multiplied by the
    diagonal inverse.  Formed by the mixed node pass itself (each row's blocks
    staged in LDS, its diagonal inverted in the same pass) they equal those of
    the separate diagonal pass bit for bit: ten batch steps give the same
    states, residuals and operands."""
    meshes, params = _build(fsm.FS_PREC_MIXED)
    params = [fsm.Params.make(**dict(p.as_dict(), subits=40)) for p in params]
    res = []
    for fused in (1, 0):
        b = fsm.Batch(meshes, params, fsm.FS_PREC_MIXED)
        b.set_tuning(node_fused=fused)
        t = b.tuning()
        assert t["gs32_scaled"] == 1 and t["node_fused"] == fused
        b.init_freestream()
        for _ in range(10):
            b.step()
        res.append([(b.get(i, "Q"), b.get(i, "R"), b.get(i, "LHS32")) for i in range(len(meshes))])
        b.close()
    for i, (a, c) in enumerate(zip(*res)):
        print("instance %d: states, residuals and operands identical %s" %
              (i, all(np.array_equal(x, y) for x, y in zip(a, c))), flush=True)
        for x, y in zip(a, c):
            assert np.array_equal(x, y), "instance %d" % i


@pytest.mark.parametrize("order,soits,limit", [(1, 0, 0), (3, 2, 1)])
def test_fp16_operands_are_the_fp32_operands_rounded(order, soits, limit):
    """FS_PREC_MIXED16 keeps FS_PREC_MIXED's arithmetic and stores the
    pre-multiplied off-diagonal sweep operands rounded to FP16: from the same
    state, a step of each gives the same FP64 residual, the same FP32
    diagonal blocks, and off-diagonal operands equal to the FP32 ones rounded
    to FP16 (round to nearest even), in every phase."""
    p = _params(order=order, soits=soits, limit=limit, foits=2, subits=40)
    m = fsm.Mesh.bump(33, 17, 0.2, 7)
    m.prepare(p)
    a = fsm.Solver(m, p, fsm.FS_PREC_MIXED)
    h = fsm.Solver(m, p, fsm.FS_PREC_MIXED16)
    a.init_freestream()
    h.init_freestream()
    ia, iau = m.array("ia"), m.array("iau")
    diag = np.zeros(ia[-1], dtype=bool)
    diag[iau] = True
    for ts in range(1, 7):
        h.set("Q", a.get("Q"))
        a.solve(ts)
        h.solve(ts)
        assert np.array_equal(a.get("R"), h.get("R"))
        la = a.get("LHS32").reshape(-1, 16)
        lh = h.get("LHS32").reshape(-1, 16)
        assert np.array_equal(la[diag], lh[diag])
        off32 = la[~diag].astype(np.float32)
        want = off32.astype(np.float16).astype(np.float64)
        nz = np.count_nonzero(off32)
        rel = np.max(np.abs(lh[~diag] - off32)) / np.max(np.abs(off32))
        print("ts %d: %d off-diagonal operands, %d non-zero, largest |operand| %.3e, max FP16 rounding %.2e of it" %
              (ts, off32.size, nz, np.max(np.abs(off32)), rel), flush=True)
        assert nz > 0
        assert np.array_equal(lh[~diag], want)
    a.close()
    h.close()


@pytest.mark.parametrize("precision,subits", [(fsm.FS_PREC_F64, 12), (fsm.FS_PREC_MIXED, 12),
                                              (fsm.FS_PREC_MIXED, 40), (fsm.FS_PREC_MIXED16, 40)])
def test_sweep_gather_forms_equal(precision, subits):
    """The 4-lane sweeps read a neighbour's dQ either whole in every lane of
    the row's quad or one component per lane, exchanged within the quad, and
    issue a row's first 4 (or 6, 8) block loads before the arithmetic: the
    same values in the same arithmetic, so the same states bit for bit (FP64;
    stored-inverse and pre-multiplied forms, FP32 and FP16 operands).  The
    FP64 path's one-thread-per-row sweep gives the same bits as well."""
    meshes, params = _build(precision)
    params = [fsm.Params.make(**dict(p.as_dict(), subits=subits)) for p in params]
    res = []
    forms = [(0, 4), (1, 4), (1, 8)] + ([(-1, 6), (0, 0)] if precision == fsm.FS_PREC_F64 else [])
    for dpp, pre in forms:
        b = fsm.Batch(meshes, params, precision)
        if dpp < 0:
            b.set_tuning(gs_lanes=1)
        else:
            b.set_tuning(gs32_dpp=dpp, gs32_pre=pre)
        b.init_freestream()
        for _ in range(8):
            b.step()
        res.append([(b.get(i, "Q"), b.get(i, "dQ")) for i in range(len(meshes))])
        b.close()
    for other in res[1:]:
        for i, (a, c) in enumerate(zip(res[0], other)):
            for x, y in zip(a, c):
                assert np.array_equal(x, y), "instance %d" % i


@pytest.mark.parametrize("precision", [fsm.FS_PREC_F64, fsm.FS_PREC_MIXED, fsm.FS_PREC_MIXED16])
def test_sweep_work_lists_follow_the_active_instances(precision):
    """Once at most three quarters of the nodes belong to running instances,
    each sweep launch covers only the chunks of its rows that hold running
    rows.  Pausing most instances, resuming them, and pausing again changes
    nothing: every instance still equals its solve alone bit for bit, and a
    paused instance is left untouched."""
    meshes, params = _build(precision)
    n = len(meshes)
    keep = {1, 3}   # the 97x33 channel and a k=2 channel keep running
    b = fsm.Batch(meshes, params, precision)
    b.use_graph(True)
    b.init_freestream()
    for _ in range(3):
        b.step()
    frozen = {}
    for i in range(n):
        if i not in keep:
            b.set_active(i, False)
            frozen[i] = b.get(i, "Q")
    for _ in range(4):
        b.step()
    for i in frozen:
        assert np.array_equal(b.get(i, "Q"), frozen[i]), "paused instance %d moved" % i
    for i in frozen:
        b.set_active(i, True)
    for _ in range(2):
        b.step()
    for i in range(n):
        if i not in keep:
            b.set_active(i, False)
    for _ in range(2):
        b.step()
    # instances in `keep` ran 11 steps, the others 5
    for i, (m, p) in enumerate(zip(meshes, params)):
        s = fsm.Solver(m, p, precision)
        s.init_freestream()
        for ts in range(1, (11 if i in keep else 5) + 1):
            s.solve(ts)
        same = np.array_equal(s.get("Q"), b.get(i, "Q"))
        print("instance %d (%s) equals alone: %s" % (i, "running" if i in keep else "paused", same), flush=True)
        assert same
        s.close()
    b.close()
