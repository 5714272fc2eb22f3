"""The opt-in solve controls of batches and serving that the reference does
not have: recovery from non-physical states and the limiter freeze.

Recovery: a solve that goes non-physical returns to its last checkpoint (the
exact state it held after an earlier step) with half its CFL target and a
restarted ramp, and with the limiter fallback its first restore switches the
Barth limiter on; the solves beside it follow exactly the path they follow
with recovery off.  The failing solve is a k=1 unlimited transonic channel
(100x35 bump, M 0.796: the 10% bump chokes the channel) that goes
non-physical about ten steps into its linear phase at any CFL target (the
oracle fails it at CFL 200, 100, 50 and 20 alike).

Stall test: a solve whose residual stops falling moves on to the next fallback;
the solves that do not stall are untouched.

Limiter freeze: once the limited phase has run the set number of steps the
limiter values stop changing, and before that the solve is the reference's.

First-order fallback: a channel that the limiter fallback does not save
converges at first order, to the oracle's first-order steady state.
"""
import os
import sys

import numpy as np
import pytest

import fsutil as U

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "flow-solver_amd"))
import workloads as W  # noqa: E402

fsm = U.fsm
pytestmark = pytest.mark.gpu


def _params(**kw):
    base = dict(nsteps=0, iter=1, jufreq=1, CFL=200.0, ramp=30, gamma=1.4, alpha=0.0, mach=0.5, subits=40,
                method=1, order=1, foits=10, soits=0, limit=0, reorder=1)
    base.update(kw)
    return fsm.Params.make(**base)


def _meshes():
    choked = W.batch_cases(256)[21]
    assert choked["kind"] == "bump" and choked["dims"] == (100, 35) and choked["k"] == 1
    cases = [
        (W.make_mesh(choked), _params(mach=choked["mach"], alpha=choked["alpha"])),
        (fsm.Mesh.bump(33, 17, 0.2, 1), _params(mach=0.5, alpha=0.5)),
        (fsm.Mesh.bump(41, 13, 0.2, 2), _params(mach=0.6, alpha=1.0, order=3, foits=5, soits=4, limit=1)),
    ]
    for m, p in cases:
        m.prepare(p)
    return [m for m, _ in cases], [p for _, p in cases]


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED, fsm.FS_PREC_F64])
def test_recovery_restores_checkpoint_and_leaves_healthy_solves_alone(precision):
    meshes, params = _meshes()
    nsteps = 40
    plain = fsm.Batch(meshes, params, precision)
    plain.init_freestream()
    r2_plain = []
    for _ in range(nsteps):
        r2, _ = plain.step()
        r2_plain.append(r2.copy())
    st0 = plain.status(0)
    print("without recovery: choked solve status %d after %d steps" % (st0[0], st0[1]), flush=True)
    assert st0[0] == fsm.INST_FAILED
    q_plain = [plain.get(i, "Q") for i in (1, 2)]
    plain.close()

    b = fsm.Batch(meshes, params, precision)
    b.init_freestream()
    b.set_recovery(2, 5)
    seen, restores = [], 0
    for k in range(nsteps):
        r2, _ = b.step()
        # the healthy solves: bit for bit the path without recovery
        assert r2[1] == r2_plain[k][1] and r2[2] == r2_plain[k][2], "step %d" % (k + 1)
        if b.status(0)[0] == fsm.INST_FAILED:
            continue
        q = b.get(0, "Q")
        n = b.recoveries(0)[0]
        if n > restores:
            restores = n
            match = [j + 1 for j, s in enumerate(seen) if np.array_equal(s, q)]
            print("step %d: restore %d returned the state held after step(s) %s" % (k + 1, n, match), flush=True)
            assert match, "a restore must return a state the solve held before"
        seen.append(q)
    st = b.status(0)
    n, fb = b.recoveries(0)
    print("with recovery: choked solve status %d, %d steps, %d restores, fallbacks %d" % (st[0], st[1], n, fb),
          flush=True)
    assert n >= 1 and not fb
    assert st[1] == nsteps or st[0] == fsm.INST_FAILED
    if st[0] == fsm.INST_FAILED:
        assert n == 2
    for i, q in zip((1, 2), q_plain):
        assert np.array_equal(b.get(i, "Q"), q)
        assert b.recoveries(i) == (0, 0)
    b.close()


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED, fsm.FS_PREC_F64])
def test_limiter_fallback_switches_the_limiter_on_at_the_first_restore(precision):
    meshes, params = _meshes()
    b = fsm.Batch(meshes[:2], params[:2], precision)
    b.init_freestream()
    b.set_recovery(3, 5, 20)
    first, seen = None, []
    for k in range(60):
        b.step()
        n, fb = b.recoveries(0)
        if n >= 1 and first is None:
            first = k + 1
            # the fallback comes with the first restore and only then, and
            # restarts the linear phase: the state after the last first-order step
            assert fb == fsm.FALLBACK_LIMITER
            assert np.array_equal(b.get(0, "Q"), seen[params[0].foits - 1])
            phi = b.get(0, "phi")
            assert np.all(phi == 1.0), "the restored solve has not run its limiter yet"
        if first is None:
            assert not fb
            seen.append(b.get(0, "Q"))
    assert first is not None, "the choked solve never failed"
    phi = b.get(0, "phi")
    st = b.status(0)
    print("fallback at step %d; status %d after %d steps, %d restores; min phi %.3f" %
          (first, st[0], st[1], b.recoveries(0)[0], phi.min()), flush=True)
    if st[0] != fsm.INST_FAILED:
        assert phi.min() < 1.0, "the limiter never acted after the fallback"
    # the healthy solve never needed it
    assert b.recoveries(1) == (0, 0)
    b.close()


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED, fsm.FS_PREC_F64])
def test_limiter_freeze_keeps_phi_after_the_set_steps(precision):
    p = _params(CFL=50.0, ramp=10, subits=16, foits=3, limit=1, mach=0.55, alpha=1.0)
    m = fsm.Mesh.bump(33, 17, 0.2, 3)
    m.prepare(p)
    freeze, nsteps = 4, 12
    ref = fsm.Batch([m], [p], precision)
    ref.init_freestream()
    fz = fsm.Batch([m], [p], precision)
    fz.init_freestream()
    fz.set_limiter_freeze(freeze)
    phis = []
    for ts in range(1, nsteps + 1):
        ref.step()
        fz.step()
        phis.append(fz.get(0, "phi"))
        if ts <= p.foits + freeze:
            # the limiter still runs: the reference's step
            assert np.array_equal(fz.get(0, "Q"), ref.get(0, "Q")), "step %d" % ts
    last_live = p.foits + freeze  # last step whose limiter pass ran
    assert all(np.all(phis[t - 1] == 1.0) for t in range(1, p.foits + 1))
    live = [phis[t - 1] for t in range(p.foits + 1, last_live + 1)]
    assert any(np.any(ph < 1.0) for ph in live), "the limiter never acted"
    assert any(not np.array_equal(live[0], ph) for ph in live[1:]), "phi did not change while the limiter ran"
    for t in range(last_live + 1, nsteps + 1):
        assert np.array_equal(phis[t - 1], phis[last_live - 1]), "phi changed after the freeze (step %d)" % t
    ref.close()
    fz.close()


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED, fsm.FS_PREC_F64])
def test_first_order_fallback_reaches_the_first_order_steady_state(precision):
    """A transonic channel (151x51 bump, M 0.798) that fails unlimited and
    again with the limiter fallback: the first-order fallback restarts its
    linear phase with nodal face states, and the solve converges to the
    first-order steady state, the oracle's (run first order from the
    freestream) within 1e-8 in density."""
    c = W.batch_cases(256)[50]
    assert c["kind"] == "bump" and c["dims"] == (151, 51) and c["k"] == 1
    p = _params(mach=c["mach"], alpha=c["alpha"])
    m = W.make_mesh(c)
    m.prepare(p)
    b = fsm.Batch([m], [p], precision)
    b.init_freestream()
    b.set_recovery(4, 10, 50, True)
    steps, drop, status = b.solve_steady(2500, 1e-10)
    n, fb = b.recoveries(0)
    print("status %d after %d steps, drop %.2e, %d restores, fallbacks %d" % (status[0], steps[0], drop[0], n, fb),
          flush=True)
    assert status[0] == fsm.INST_CONVERGED
    assert fb & fsm.FALLBACK_FIRST_ORDER
    rho = b.get(0, "Q")[0:4 * m.info()["nn"]:4]
    b.close()
    p1 = _params(mach=c["mach"], alpha=c["alpha"], foits=1 << 28)
    o = U.Oracle(m, p1)
    o.init_freestream()
    r0 = None
    for ts in range(1, 2001):
        o.solve(ts, multicolor=1)
        r = np.sqrt(o.s.last_rms)
        if ts == 11:
            r0 = r
        if r0 and r / r0 <= 1e-11:
            break
    rho_o = o["Q"][0:4 * m.info()["nn"]:4]
    diff = np.linalg.norm(rho - rho_o) / np.linalg.norm(rho_o)
    print("oracle first order: %d steps; density L2 diff %.2e" % (ts, diff), flush=True)
    assert r / r0 <= 1e-11
    assert diff <= 1e-8


@pytest.mark.parametrize("precision", [fsm.FS_PREC_MIXED16, fsm.FS_PREC_MIXED])
def test_stall_starts_the_next_fallback(precision):
    """A transonic channel (180x60 bump, M 0.745, a request of the serving
    stream) fails unlimited and restarts with the limiter fallback.  With
    FP16 sweep operands its frozen-limiter iteration then cycles near a 1e-1
    drop and reaches the step limit; with the stall test (2 sqrt(nodes) steps
    without halving the residual) the stall starts the first-order fallback
    and the solve converges.  With FP32 operands the frozen limiter takes
    other values (the frozen state depends on the iterate path), the solve
    converges without stalling, and the stall test changes nothing.  The
    solve beside it never stalls and follows exactly the path it follows with
    the test off."""
    c = W.serving_cases(287)[139]
    assert c["kind"] == "bump" and c["dims"] == (180, 60) and c["k"] == 1
    cases = [(W.make_mesh(c), _params(mach=c["mach"], alpha=c["alpha"])),
             (fsm.Mesh.bump(33, 17, 0.2, 1), _params(mach=0.5, alpha=0.5))]
    for m, p in cases:
        m.prepare(p)
    out = {}
    for stall in (0.0, 2.0):
        b = fsm.Batch([m for m, _ in cases], [p for _, p in cases], precision)
        b.init_freestream()
        b.set_recovery(4, 10, 50, True)
        b.set_stall(stall)
        steps, drop, status = b.solve_steady(2000, 1e-10)
        out[stall] = (steps, drop, status, [b.recoveries(i) for i in range(2)], [b.get(i, "Q") for i in range(2)])
        print("stall %.0f: status %s steps %s drop %s recoveries %s" % (stall, status, steps, drop, out[stall][3]),
              flush=True)
        b.close()
    steps, drop, status, rec, q = out[0.0]
    steps_s, drop_s, status_s, rec_s, q_s = out[2.0]
    assert status[1] == fsm.INST_CONVERGED and rec[1] == (0, 0)
    assert status_s[1] == fsm.INST_CONVERGED and rec_s[1] == (0, 0) and steps_s[1] == steps[1]
    assert np.array_equal(q_s[1], q[1])
    assert rec[0][1] == fsm.FALLBACK_LIMITER
    if precision == fsm.FS_PREC_MIXED16:
        assert status[0] == fsm.INST_MAX_STEPS
        assert status_s[0] == fsm.INST_CONVERGED
        assert rec_s[0][1] == fsm.FALLBACK_LIMITER | fsm.FALLBACK_FIRST_ORDER | fsm.FALLBACK_STALL
    else:
        assert status[0] == fsm.INST_CONVERGED
        assert status_s[0] == fsm.INST_CONVERGED and rec_s[0] == rec[0] and steps_s[0] == steps[0]
        assert np.array_equal(q_s[0], q[0])
