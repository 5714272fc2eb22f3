"""C-ABI checks that need no GPU: every entry point declared in
include/fs_capi.h is exported by the built library, the device entry points
fail with a status (not a crash, not a CPU fallback) when no HIP device is
present, and the batch's color-launch schedule gives every instance exactly
the sweep order it has when solved alone."""
import ctypes as C
import os
import re

import pytest

import fsutil as U

fsm = U.fsm
HEADER = os.path.join(U.ROOT, "include", "fs_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(fs_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 40, names
    lib = C.CDLL(fsm.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _no_device():
    return fsm.device_count() == 0


@pytest.mark.skipif(not _no_device(), reason="a HIP device is present")
def test_device_entry_points_fail_without_device():
    p = fsm.Params.make(order=1, subits=4)
    m = fsm.Mesh.bump(9, 5, 0.2, 1).prepare(p)
    with pytest.raises(fsm.FlowSolverError) as e:
        fsm.Solver(m, p)
    assert e.value.code == -4
    with pytest.raises(fsm.FlowSolverError) as e:
        fsm.Batch([m, m], p)
    assert e.value.code == -4


def test_footprint_grows_with_mesh():
    p = fsm.Params.make(order=3, subits=4)
    small = fsm.Mesh.bump(17, 9, 0.2, 1).prepare(p)
    big = fsm.Mesh.bump(33, 17, 0.2, 1).prepare(p)
    fs, fb = small.footprint(fsm.FS_PREC_F64), big.footprint(fsm.FS_PREC_F64)
    assert 0 < fs < fb
    # FP32 blocks halve the largest array
    assert big.footprint(fsm.FS_PREC_MIXED) < fb
    nn, nnz = big.info()["nn"], big.info()["nnz"]
    assert fb >= nnz * 128 + nn * 128


def _schedule(ncolor, subits):
    """(color, backward) launches of the device sweep order."""
    out, last = [], -1
    for it in range(subits):
        for cc in range(ncolor):
            c = cc if it % 2 == 0 else ncolor - 1 - cc
            if c != last:
                out.append((c, it % 2))
                last = c
    return out


def _visits(ncolor_inst, ncolor_max, subits):
    """Colors instance sees inside a batch: empty colors are no-ops and a
    backward launch of the instance's last color skips it."""
    seen = []
    for c, bwd in _schedule(ncolor_max, subits):
        if c >= ncolor_inst or (bwd and c == ncolor_inst - 1):
            continue
        seen.append(c)
    return seen


@pytest.mark.parametrize("ncmax", [1, 2, 3, 4, 5, 7])
def test_batch_schedule_equals_alone(ncmax):
    for nci in range(1, ncmax + 1):
        for subits in range(0, 13):
            alone = [c for c, _ in _schedule(nci, subits)]
            assert _visits(nci, ncmax, subits) == alone, (nci, ncmax, subits)


def test_schedule_drops_turnaround_repeat():
    # every sweep after the first starts at the color the previous one ended
    # on; that launch is dropped
    s = _schedule(4, 3)
    assert [c for c, _ in s] == [0, 1, 2, 3, 2, 1, 0, 1, 2, 3]


@pytest.mark.parametrize("kw,frag", [(dict(order=2), "order"), (dict(order=4), "order"), (dict(limit=3), "limit"),
                                     (dict(method=0), "method"), (dict(unsteady=1), "unsteady"),
                                     (dict(jufreq=0), "jufreq"), (dict(gamma=1.0), "gamma")])
def test_unsupported_params_rejected(kw, frag):
    """Run controls outside the path are rejected at creation with
    FS_ERR_UNSUPPORTED (before any device work), naming the field."""
    base = dict(order=1, foits=2, limit=1, reorder=1, mach=0.5)
    m = fsm.Mesh.bump(9, 5, 0.1, 1)
    m.prepare(fsm.Params.make(**base))
    base.update(kw)
    with pytest.raises(fsm.FlowSolverError) as e:
        fsm.Solver(m, fsm.Params.make(**base))
    assert e.value.code == -7, e.value
    assert frag in str(e.value)


def test_batch_rejects_mixed_sweep_controls():
    """Instances of one batch share subits, iter and jufreq (one launch per
    sweep for all of them)."""
    m1, m2 = fsm.Mesh.bump(9, 5, 0.1, 1), fsm.Mesh.bump(11, 5, 0.1, 2)
    p1 = fsm.Params.make(order=1, foits=2, limit=1, reorder=1, mach=0.5, subits=10)
    p2 = fsm.Params.make(order=1, foits=2, limit=1, reorder=1, mach=0.5, subits=12)
    m1.prepare(p1)
    m2.prepare(p2)
    with pytest.raises(fsm.FlowSolverError) as e:
        fsm.Batch([m1, m2], [p1, p2])
    assert e.value.code == -1 and "subits" in str(e.value)


def test_mesh_from_arrays_validates_input():
    """fs_mesh_create (the grid file's contents from arrays) rejects node
    indices out of range in triangles, quadrilaterals and segments, and
    a mesh with both element types prepares."""
    x = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]
    tri = [(0, 1, 4), (0, 4, 3)]
    quad = [(1, 2, 5, 4)]
    seg = [(0, 1), (1, 2), (2, 5), (5, 4), (4, 3)]
    seg_closed = seg + [(3, 0)]
    m = fsm.Mesh.from_arrays(x, tri, quad, [0, 6], seg_closed, [1])
    m.prepare(fsm.Params.make(order=1, subits=4))
    info = m.info()
    assert (info["ntri"], info["nquad"], info["nsub"]) == (2, 1, 10)
    for bad_tri, bad_quad, bad_seg in (([(0, 1, 9)], quad, seg), (tri, [(1, 2, 5, 6)], seg), (tri, quad, [(0, 7)])):
        with pytest.raises(fsm.FlowSolverError):
            fsm.Mesh.from_arrays(x, bad_tri, bad_quad, [0, len(bad_seg)], bad_seg, [1])


def test_solve_controls_reject_bad_arguments():
    """The stall test, the k=2 deferral switch and the server's stall field
    check their arguments before touching a device."""
    L = fsm.lib()
    assert L.fs_batch_set_stall(None, 2.0) == -1
    assert L.fs_batch_defer_k2(None, 1) == -1
    for bad in (-1.0, float("nan")):
        with pytest.raises(fsm.FlowSolverError) as e:
            fsm.Server(stall_steps_per_sqrt_node=bad)
        assert e.value.code == -1
