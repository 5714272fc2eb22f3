"""fs_state_norms_noalloc (the entry point that allocates nothing) against
fs_state_norms (the same sums with per-call partials) through the C-ABI:
bit for bit on two calls in a row, the second on a smaller input, plus the
size limit and a numpy check of the sums."""
import ctypes as C

import numpy as np
import pytest

import fsutil as U

fsm = U.fsm
pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def to_device(a):
    d = C.c_void_p()
    assert hip().hipMalloc(C.byref(d), a.nbytes) == 0
    assert hip().hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
    return d


def from_device(d, n):
    out = np.zeros(n)
    assert hip().hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0
    return out


def test_noalloc_norms_match_allocating_variant():
    lib = fsm.lib()
    rng = np.random.default_rng(12)
    outs = [to_device(np.full(2, np.nan)), to_device(np.full(2, np.nan))]
    try:
        # the second call is smaller: the shared partials still hold the
        # first call's chunks beyond its range
        for nn in (300_001, 5_000):
            R = rng.standard_normal(4 * nn)
            dQ = 1e-3 * rng.standard_normal(4 * nn)
            dR, ddQ = to_device(R), to_device(dQ)
            try:
                print("norms: %d nodes" % nn, flush=True)
                assert lib.fs_state_norms(nn, dR, ddQ, outs[0], None) == 0, lib.fs_last_error()
                assert lib.fs_state_norms_noalloc(nn, dR, ddQ, outs[1], None) == 0, lib.fs_last_error()
                assert hip().hipDeviceSynchronize() == 0
                a, b = from_device(outs[0], 2), from_device(outs[1], 2)
                print("  allocating %r  no-alloc %r" % (a.tolist(), b.tolist()), flush=True)
                np.testing.assert_array_equal(a, b)
                ref = np.array([np.sum(R * R), np.sum(dQ * dQ)])
                assert np.all(np.abs(a - ref) <= 1e-12 * ref)
            finally:
                hip().hipFree(dR)
                hip().hipFree(ddQ)
        # beyond the fixed buffer: rejected on the host, nothing launched
        rc = lib.fs_state_norms_noalloc(65536 * 1024 + 1, outs[0], outs[0], outs[1], None)
        assert rc == -1 and b"at most" in lib.fs_last_error()
    finally:
        for d in outs:
            hip().hipFree(d)


def test_noalloc_norms_equal_solver_norms():
    """The sums of a solver's own R and dQ equal the r2 fs_solve reports."""
    lib = fsm.lib()
    p = fsm.Params.make(nsteps=5, iter=1, jufreq=1, CFL=20.0, ramp=10, gamma=1.4, alpha=1.0, mach=0.5, subits=10,
                        method=1, order=1, foits=2, limit=1, reorder=1)
    m = fsm.Mesh.bump(65, 33, 0.2, 3)
    m.prepare(p)
    s = fsm.Solver(m, p, fsm.FS_PREC_F64)
    s.init_freestream()
    for ts in range(1, 4):
        dq, r2 = s.solve(ts)
    nn = m.info()["nn"]
    R, dQ = s.get("R"), s.get("dQ")
    dR, ddQ, dout = to_device(R), to_device(dQ), to_device(np.zeros(2))
    try:
        assert lib.fs_state_norms_noalloc(nn, dR, ddQ, dout, None) == 0, lib.fs_last_error()
        assert hip().hipDeviceSynchronize() == 0
        out = from_device(dout, 2)
    finally:
        for d in (dR, ddQ, dout):
            hip().hipFree(d)
    print("solver r2 %r dq %r; no-alloc %r" % (r2, dq, out.tolist()), flush=True)
    assert out[0] == r2
    assert np.sqrt(out[1]) == dq
