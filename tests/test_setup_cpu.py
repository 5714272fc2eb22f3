"""Host preprocessing against the reference build's golden vectors (bit-exact
integer maps, renumbering and stencils; bit-exact FP64 geometry) plus
structural invariants that a dropped term or wrong index would break."""
import os

import numpy as np
import pytest

import fsutil as U

fsm = U.fsm
CASES = ["bump17x9_o1", "bump17x9_o3", "cfg1_97x33", "airfoil32x9_o1", "bumpmix17x9_o1"]


def _load(name):
    return np.load(os.path.join(U.GOLDEN, name + ".npz"))


def _mesh_for(g):
    nx, ny, seed, order, foits, soits, limit, nsteps = (int(v) for v in g["params"])
    cfl, mach, alpha, ramp, subits = g["params_f"]
    p = fsm.Params.make(order=order, foits=foits, soits=soits, limit=limit, mach=mach, alpha=alpha, CFL=cfl,
                        ramp=int(ramp), subits=int(subits), method=1, iter=1, jufreq=1, reorder=1, gamma=1.4)
    m = U.golden_mesh(g)
    m.prepare(p)
    return p, m


@pytest.mark.parametrize("case", CASES)
def test_maps_bitexact(case):
    g = _load(case)
    p, m = _mesh_for(g)
    nn = m.info()["nn"]
    np.testing.assert_array_equal(m.array("nn_map"), g["s_nn_map"][1:] - 1)
    np.testing.assert_array_equal(m.array("nnsn"), g["s_nnsn"][1:nn + 2])
    np.testing.assert_array_equal(m.array("nsn"), g["s_nsn"] - 1)
    np.testing.assert_array_equal(m.array("esn"), g["s_esn"] - 1)
    np.testing.assert_array_equal(m.array("edges"), g["s_edges"][2:] - 1)
    np.testing.assert_array_equal(m.array("csn"), g["s_csn"] - 1)
    np.testing.assert_array_equal(m.array("tri"), g["s_c2n_tri"][3:] - 1)
    if "s_c2n_quad" in g.files:
        np.testing.assert_array_equal(m.array("quad"), g["s_c2n_quad"][4:] - 1)
    se = g["s_subedges"].reshape(-1, 2)[1:]
    np.testing.assert_array_equal(m.array("sub_edge"), se[:, 0] - 1)
    np.testing.assert_array_equal(m.array("sub_elem"), se[:, 1] - 1)
    np.testing.assert_array_equal(m.array("bedge").reshape(-1, 5), g["s_bedges"].reshape(-1, 5)[1:] - 1)
    np.testing.assert_array_equal(m.array("ia"), g["s_ia"][1:nn + 2])
    np.testing.assert_array_equal(m.array("iau"), g["s_iau"][1:])
    np.testing.assert_array_equal(m.array("ja"), g["s_ja"] - 1)


@pytest.mark.parametrize("case", CASES)
def test_geometry_bitexact(case):
    g = _load(case)
    p, m = _mesh_for(g)
    pairs = [("x", "x", 2), ("sub_n", "xn_subedges", 3), ("sub_xm", "xm_subedges", 2), ("sub_xg", "xg_subedges", 6),
             ("el_cent", "el_cent", 2), ("bedge_n", "xn_bedges", 3), ("bedge_xg", "xg_bedges", 6),
             ("cv_area", "cv_area", 1), ("moments", "Moments", 9), ("edge_weight", "edge_weight", 4)]
    for mine, ref, skip in pairs:
        a, b = m.array(mine), g["s_" + ref][skip:]
        assert a.shape == b.shape, mine
        assert np.array_equal(a, b), (mine, np.max(np.abs(a - b)))


@pytest.mark.parametrize("case", ["bump17x9_o3", "cfg1_97x33", "bumpmix17x9_o3"])
def test_k2_stencils_bitexact(case):
    g = _load(case)
    p, m = _mesh_for(g)
    nn = m.info()["nn"]
    g1, gn, gs = g["s_gnnsn1"], g["s_gnnsn"], g["s_gnsn"]
    st, off = [], [0]
    for i in range(1, nn + 1):
        a, b = gn[g1[i] + 1], gn[g1[i + 1]]
        st.extend(gs[a:b] - 1)
        off.append(len(st))
    np.testing.assert_array_equal(m.array("st_off"), np.array(off, dtype=np.int32))
    np.testing.assert_array_equal(m.array("st_nodes"), np.array(st, dtype=np.int32))


def test_invariants_bump():
    p = fsm.Params.make(order=3, foits=1, soits=1, limit=1, reorder=1)
    m = fsm.Mesh.bump(41, 15, 0.2, 9)
    m.prepare(p)
    info = m.info()
    nn, ne, ntri = info["nn"], info["ne"], info["ntri"]
    assert nn - ne + ntri == 1  # Euler characteristic of a disk
    edges = m.array("edges").reshape(-1, 2)
    assert np.all(edges[:, 0] < edges[:, 1])
    # every edge hit by one or two subedges
    hits = np.bincount(m.array("sub_edge"), minlength=ne)
    assert hits.min() >= 1 and hits.max() <= 2
    # total control-volume area equals the domain area: channel 3x1 minus the bump segment
    R, h = 1.3, 0.1
    half = np.arcsin(0.5 / R)
    seg = R * R * half - 0.5 * 1.0 * (R - h)
    # the bump is a polyline at the mesh's resolution; use the exact polygon area instead
    x = m.array("x")[:2 * nn].reshape(-1, 2)
    bseg = m.array("bseg").reshape(-1, 2)
    poly = 0.5 * np.sum(x[bseg[:, 0], 0] * x[bseg[:, 1], 1] - x[bseg[:, 1], 0] * x[bseg[:, 0], 1])
    assert abs(np.sum(m.array("cv_area")) - poly) < 1e-12 * poly
    assert abs(poly - (3.0 - seg)) < 1e-3
    # closure: sum of n*len around each control volume vanishes
    sn = m.array("sub_n").reshape(-1, 3)
    se = m.array("sub_edge")
    acc = np.zeros((nn, 2))
    L, Rr = edges[se, 0], edges[se, 1]
    np.add.at(acc, L, sn[:, :2] * sn[:, 2:3])
    np.add.at(acc, Rr, -sn[:, :2] * sn[:, 2:3])
    be = m.array("bedge").reshape(-1, 5)
    bn = m.array("bedge_n").reshape(-1, 3)
    np.add.at(acc, be[:, 0], bn[:, :2] * bn[:, 2:3])
    assert np.max(np.abs(acc)) < 1e-13
    # moments of a node about itself: first moments equal the CV centroid offset; M2,M3 >= 0
    mo = m.array("moments").reshape(-1, 9)
    assert np.all(mo[:, 2] > 0) and np.all(mo[:, 3] > 0)
    # coloring is a proper distance-1 coloring of the matrix graph
    col = m.array("color")
    assert np.all(col[edges[:, 0]] != col[edges[:, 1]])
    co, rows = m.array("color_off"), m.array("color_rows")
    assert co[-1] == nn and np.array_equal(np.sort(rows), np.arange(nn))
    # block CSR: diagonal last in each row, transposed slots consistent
    ia, iau, ja = m.array("ia"), m.array("iau"), m.array("ja")
    assert np.all(iau == ia[1:] - 1) and np.all(ja[iau] == np.arange(nn))


def test_grid_format_roundtrip(tmp_path):
    m = fsm.Mesh.bump(13, 7, 0.2, 4)
    g, b = str(tmp_path / "a.mesh"), str(tmp_path / "bc.dat")
    m.write(g, b)
    m2 = fsm.Mesh.read(g, b)
    assert m2.info()["nn"] == 91 and m2.info()["ntri"] == 144
    np.testing.assert_array_equal(m.array("x"), m2.array("x"))
    np.testing.assert_array_equal(m.array("tri"), m2.array("tri"))
    np.testing.assert_array_equal(m.array("bseg"), m2.array("bseg"))
    np.testing.assert_array_equal(m2.array("bbc"), [1, 0, 1, 0])
    p = fsm.Params.make(order=3, CFL=12.5, mach=0.675, alpha=1.25, foits=7)
    fsm.write_params(str(tmp_path / "euler.params"), p)
    q = fsm.read_params(str(tmp_path / "euler.params"))
    assert q.as_dict() == p.as_dict()
    # malformed inputs are rejected with a status, not an exit
    (tmp_path / "bad.mesh").write_text("#nodes\n3\n0 0\n")
    with pytest.raises(fsm.FlowSolverError):
        fsm.Mesh.read(str(tmp_path / "bad.mesh"), None)


def test_airfoil_omesh():
    p = fsm.Params.make(order=3, foits=1, soits=1, reorder=1)
    m = fsm.Mesh.airfoil(65, 40, 1e-3, 30.0, 0.1, 3)
    m.prepare(p)
    info = m.info()
    assert info["nn"] - info["ne"] + info["ntri"] == 0  # annulus
    assert np.all(m.array("cv_area") > 0)
    x = m.array("x")[:2 * info["nn"]].reshape(-1, 2)
    assert np.max(np.hypot(x[:, 0] - 0.5, x[:, 1])) <= 30.0 + 1e-9


def test_restart_file_roundtrip(tmp_path):
    """A restart file (write_mesh's layout, mesh_io.C:318-395, plus the ghost
    extension) read back with isrestart = 1 gives the same preprocessed mesh
    and the same state, bit for bit (%.17e keeps every double)."""
    p = fsm.Params.make(order=3, foits=2, soits=2, reorder=1)
    m = fsm.Mesh.bump(21, 9, 0.2, 4).prepare(p)
    info = m.info()
    rng = np.random.default_rng(1)
    Q = rng.uniform(0.5, 2.0, 4 * (info["nn"] + info["ng"]))
    path = str(tmp_path / "restart.mesh")
    bc = str(tmp_path / "bc.dat")
    with open(bc, "w") as f:
        f.write("".join("%d\n" % v for v in m.array("bbc")))
    m.write_restart(path, Q, 37, 12.5, ghosts=True)
    r = fsm.Mesh.read(path, bc)
    assert r.restart_info() == (True, 37, 12.5, True)
    pr = fsm.Params.make(order=3, foits=2, soits=2, reorder=1, isrestart=1)
    r.prepare(pr)
    for name in ("x", "edges", "sub_n", "bedge", "ia", "ja", "color", "color_rows", "st_nodes", "st_W", "cv_area"):
        np.testing.assert_array_equal(r.array(name), m.array(name), err_msg=name)
    # the reference's layout: constants, then one line of 4 values per real node
    lines = open(path).read().splitlines()
    k = lines.index("#Number of Constants")
    assert lines[k + 1:k + 3] == ["2", "37"] and float(lines[k + 3]) == 12.5
    assert lines[k + 4] == "#Number of Variables" and lines[k + 5] == "4"
    vals = np.array([float(v) for v in " ".join(lines[k + 6:k + 6 + info["nn"]]).split()])
    np.testing.assert_array_equal(vals, Q[:4 * info["nn"]])
    # without the extension the file ends after the real-node states
    m.write_restart(path, Q, 37, 12.5, ghosts=False)
    assert len(open(path).read().splitlines()) == k + 6 + info["nn"]
    assert fsm.Mesh.read(path, bc).restart_info() == (True, 37, 12.5, False)
