"""Accessibility grids on the host: the oracle's restatement pinned by the
reference's own known answers, pm_block_pockets (host code of libpmhip.so)
against the oracle, the Grid helpers and the .cube format.

Known answers (test/grid.jl):
  :17       required_n_pts(Box(1, 10, 5, 90 deg ...), 0.1) == (11, 101, 51)
  :19-24    write_cube -> read_cube round trip (atol 1e-5)
  :27-40    _arg_nearest_neighbor / _apply_pbc_to_index cases
  :42-99    LTA and SOD, CH4 / UFF, resolution 2 A, energy_tol 0 kJ/mol:
            LTA blocks 8 segments, SOD blocks every segment, porosity drops,
            no-blocking porosity == porosity before blocking
  :102-125  LTA replicated (2, 3, 1): the unit cell's grid is the corner of
            the replicated one; accessible(rep grid, xf) ==
            accessible(grid, xf, (2, 3, 1))
  :127-146  SBMOF-1 and CAXVII_clean at 5 kJ/mol: no segment blocked
  :149-153  xf_to_id cases
The energy grids of these cases come from the oracle's FP64 C code.  The
oracle's cycle search is checked against brute force on small random graphs.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def oacc():
    from oracle import accessibility
    return accessibility


@pytest.fixture(scope="module")
def uff(pm):
    return pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))


def oracle_grid_kj(pm, orc, x, ff, n_pts):
    """CH4 energy grid of crystal x (kJ/mol, (n1, n2, n3)) from the oracle."""
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], None, ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), np.zeros((1, 3)), [ff.index("CH4")], np.zeros((0, 3)),
                         np.zeros(0))
    return o.energy_grid(n_pts).reshape(n_pts, order="F") * 8.314 / 1000.0


# ------------------------------------------------------------ indexing ----
def test_nearest_neighbor_kats(pm, oacc):
    from porousmaterials_jl_amd.accessibility import _apply_pbc_to_index, _arg_nearest_neighbor
    cases = [((10, 20, 30), [0.001, 0.001, 0.001], [1, 1, 1]),
             ((10, 20, 30), [0.999, 0.999, 0.999], [10, 20, 30]),
             ((3, 3, 3), [0.001, 0.001, 0.001], [1, 1, 1]),
             ((3, 3, 3), [0.999, 0.999, 0.999], [3, 3, 3]),
             ((3, 3, 3), [0.001, 0.001, 0.24], [1, 1, 1]),
             ((3, 3, 3), [0.001, 0.001, 0.26], [1, 1, 2]),
             ((3, 3, 3), [0.74, 0.001, 0.26], [2, 1, 2]),
             ((3, 3, 3), [0.76, 0.001, 0.26], [3, 1, 2]),
             ((4, 6, 11), [0.32, 0.42, 0.61], [2, 3, 7])]
    for n, xf, want in cases:
        assert _arg_nearest_neighbor(n, xf) == want
        assert oacc.arg_nearest_neighbor(n, xf) == want
    assert _apply_pbc_to_index([0, 21, 31], (10, 20, 30)) == [10, 1, 1]
    assert oacc.apply_pbc_to_index([0, 21, 31], (10, 20, 30)) == [10, 1, 1]
    # round(Int, x) rounds half to even: 0.125 * 4 = 0.5 -> 0, 0.375 * 4 = 1.5 -> 2
    assert oacc.arg_nearest_neighbor((5, 5, 5), [0.125, 0.375, 0.625]) == [1, 3, 3]
    assert _arg_nearest_neighbor((5, 5, 5), [0.125, 0.375, 0.625]) == [1, 3, 3]


def test_grid_index_helpers(pm):
    box = pm.Box(1.0, 10.0, 5.0, math.pi / 2, math.pi / 2, math.pi / 2)
    assert pm.required_n_pts(box, 0.1) == (11, 101, 51)
    n = (4, 4, 4)
    assert pm.xf_to_id(n, [0.0001, 0.0001, 0.0001]) == [1, 1, 1]
    assert pm.xf_to_id(n, [0.9999, 0.9999, 0.9999]) == [4, 4, 4]
    assert pm.xf_to_id(n, [0.0001, 0.2400, 0.2600]) == [1, 1, 2]
    assert pm.xf_to_id(n, [-0.0001, 0.2400, 1.01]) == [4, 1, 1]
    assert pm.id_to_xf((1, 3, 5), (5, 5, 5)) == [0.0, 0.5, 1.0]


def test_accessible_wraps_like_the_reference(pm, oacc):
    """The wrap 0 -> n maps the neighbour below point 1 onto point n, the
    same place as point 1 (not point n - 1): a grid open only at point n - 1
    of an axis is not seen from xf = 0."""
    data = np.zeros((5, 5, 5), dtype=bool)
    data[3, 0, 0] = True  # point 4 of 5 along a
    g = pm.Grid(pm.unit_cube(), (5, 5, 5), data, "accessibility", np.zeros(3))
    assert not pm.accessible(g, [0.0, 0.0, 0.0])
    assert not oacc.accessible(data, [0.0, 0.0, 0.0])
    data[4, 0, 0] = True  # point 5 of 5 == fractional 1
    assert pm.accessible(g, [0.0, 0.0, 0.0]) and oacc.accessible(data, [0.0, 0.0, 0.0])
    rng = np.random.default_rng(3)
    for _ in range(200):
        d = rng.random((4, 5, 6)) < 0.1
        gg = pm.Grid(pm.unit_cube(), d.shape, d, "accessibility", np.zeros(3))
        for xf in rng.random((20, 3)):
            assert pm.accessible(gg, xf) == oacc.accessible(d, xf)
            assert pm.accessible(gg, xf, (2, 1, 3)) == oacc.accessible(d, xf, (2, 1, 3))


# ---------------------------------------------------------------- cube ----
def test_cube_round_trip(pm, tmp_path, monkeypatch):
    """test/grid.jl:19-24: write_cube then read_cube gives the grid back to
    atol 1e-5 (six decimals in the file)."""
    import porousmaterials_jl_amd.grid as gridmod
    monkeypatch.setitem(gridmod.PATHS, "grids", str(tmp_path))
    rng = np.random.default_rng(0)
    for n in ((3, 3, 3), (4, 5, 7), (2, 3, 12)):
        g = pm.Grid(pm.Box(0.7, 0.8, 0.9, 1.5, 1.6, 1.7), n, rng.random(n), "kJ_mol", np.array([1.0, 2.0, 3.0]))
        pm.write_cube(g, "test_grid", verbose=False)
        g2 = pm.read_cube("test_grid.cube")
        assert g.isapprox(g2, atol=1e-5)
        assert not g.isapprox(pm.Grid(g.box, n, g.data + 1e-3, "kJ_mol", g.origin), atol=1e-5)
    text = open(tmp_path / "test_grid.cube").read().split("\n")
    assert text[0] == "Units of data: kJ_mol" and text[1] == "Loop order: x, y, z"
    assert text[2] == "0 1.000000 2.000000 3.000000"
    # a blocked grid point written as -1, overlaps as Inf (Julia prints "Inf")
    d = np.ones((2, 2, 2))
    d[0, 0, 0] = np.inf
    pm.write_cube(pm.Grid(pm.unit_cube(), (2, 2, 2), d, "K", np.zeros(3)), "inf", verbose=False)
    g3 = pm.read_cube("inf")
    assert open(tmp_path / "inf.cube").read().split("\n")[6].startswith("Inf 1.000000e+00")
    assert np.isinf(g3.data[0, 0, 0]) and g3.data[1, 1, 1] == 1.0


# -------------------------------------------------------- cycle search ----
def _brute_force_cycles(adj):
    nv = len(adj)
    edges = {(v, w) for v in range(nv) for w in adj[v]}
    found = set()
    for length in range(1, nv + 1):
        for path in itertools.permutations(range(nv), length):
            if path[0] != min(path):
                continue
            if all((path[i], path[(i + 1) % length]) in edges for i in range(length)):
                found.add(path)
    return found


def test_oracle_cycles_vs_brute_force(oacc):
    rng = np.random.default_rng(11)
    for _ in range(60):
        n_seg = int(rng.integers(1, 4))
        conns = []
        for _ in range(int(rng.integers(0, 4))):
            s, d = int(rng.integers(1, n_seg + 1)), int(rng.integers(1, n_seg + 1))
            t = [(0, 0, 1), (1, 0, 0), (0, 1, 0)][int(rng.integers(0, 3))]
            if (s, d, t) in conns:
                continue
            conns += [(s, d, t), (d, s, (-t[0], -t[1], -t[2]))]
        adj, _ = oacc.expanded_graph(n_seg, conns)
        if len(adj) > 8:
            continue
        got = {tuple(c) for c in oacc.simple_cycles(adj)}
        assert got == _brute_force_cycles(adj)


def test_oracle_pocket_cases(oacc):
    """Hand-made grids with known answers."""
    hi, lo = 100.0, -1.0
    # a straight channel along c: one segment, accessible through its own image
    e = np.full((5, 5, 5), hi)
    e[2, 2, :] = lo
    acc, seg, n_seg, n_bl, _ = oacc.block(e, 0.0)
    assert (n_seg, n_bl) == (1, 0) and acc[2, 2, :].all() and acc.sum() == 5
    # a closed cage: blocked
    e = np.full((6, 6, 6), hi)
    e[2:4, 2:4, 2:4] = lo
    acc, _, n_seg, n_bl, n_cyc = oacc.block(e, 0.0)
    assert (n_seg, n_bl, n_cyc) == (1, 1, 0) and not acc.any()
    # a cage cut by the c face: two segments joined across the face and back
    # (net translation 0): both blocked
    e = np.full((6, 6, 6), hi)
    e[2:4, 2:4, 0] = lo
    e[2:4, 2:4, 5] = lo
    acc, _, n_seg, n_bl, n_cyc = oacc.block(e, 0.0)
    assert (n_seg, n_bl) == (2, 2) and n_cyc >= 1 and not acc.any()
    # a channel that leaves through a and re-enters from b in one segment:
    # an L in the cell whose two ends meet their own images
    e = np.full((6, 6, 6), hi)
    e[:, 0, 3] = lo
    e[0, :, 3] = lo
    acc, _, n_seg, n_bl, _ = oacc.block(e, 0.0)
    assert (n_seg, n_bl) == (1, 0) and acc.sum() == 11
    # seeds need E <= tol, neighbours E < tol: a point at exactly tol seeds a
    # segment that bridges its < tol neighbours; scanned later it is -1
    e = np.full((1, 1, 5), hi)
    e[0, 0, 0] = 0.0
    e[0, 0, 1] = lo
    seg, n_seg = oacc.segment_grid(e, 0.0)
    assert n_seg == 1 and list(seg[0, 0, :2]) == [1, 1]
    e[0, 0, 0], e[0, 0, 1] = lo, 0.0
    seg, n_seg = oacc.segment_grid(e, 0.0)
    assert n_seg == 1 and list(seg[0, 0, :2]) == [1, -1]


def _random_landscape(rng, shape, n_smooth=2):
    e = rng.standard_normal(shape)
    for _ in range(n_smooth):
        e = sum(np.roll(e, s, axis=a) for a in range(3) for s in (-1, 0, 1)) / 9.0
    return e


def test_block_pockets_library_vs_oracle(pm, oacc):
    """pm_block_pockets (host C++) equals the oracle point for point: labels,
    accessibility and the (segments, blocked, cycles) counts."""
    rng = np.random.default_rng(5)
    shapes = [(6, 6, 6), (9, 7, 5), (1, 8, 8), (12, 1, 3), (2, 2, 2), (10, 10, 10), (16, 5, 9)]
    n_cases = 0
    for shape in shapes:
        for _ in range(6):
            e = _random_landscape(rng, shape)
            tol = float(np.quantile(e, rng.uniform(0.2, 0.6)))
            acc, seg, n_seg, n_bl, n_cyc = oacc.block(e, tol)
            lacc, lseg, counts = pm.device.block_pockets(e, tol)
            assert np.array_equal(lseg, seg), shape
            assert np.array_equal(lacc, acc), shape
            assert counts == (n_seg, n_bl, n_cyc), (shape, counts, (n_seg, n_bl, n_cyc))
            nb, _, nb_counts = pm.device.block_pockets(e, tol, block_pockets=False)
            assert np.array_equal(nb, e < tol) and nb_counts == (0, 0, 0)
            n_cases += 1
    assert n_cases == 6 * len(shapes)


def test_block_pockets_errors(pm):
    from porousmaterials_jl_amd._lib import PMError
    e = _random_landscape(np.random.default_rng(2), (10, 10, 10))
    with pytest.raises(PMError, match="max_cycles"):
        pm.device.block_pockets(e, float(np.median(e)), max_cycles=1)
    with pytest.raises(PMError, match="n_pts"):
        pm.device.block_pockets(np.zeros((0, 2, 2)), 0.0)


# ------------------------------------------------ reference known answers ----
@pytest.mark.parametrize("name", ["LTA", "SOD"])
def test_zeolite_pockets_kat(pm, orc, oacc, uff, name):
    x = pm.read_cif(golden("crystals", f"{name}.cif"))
    e = oracle_grid_kj(pm, orc, x, uff, pm.required_n_pts(x.box, 2.0))
    acc, seg, n_seg, n_bl, _ = oacc.block(e, 0.0)
    assert n_bl > 0
    if name == "LTA":
        assert n_bl == 8
    else:
        assert not acc.any()
    assert np.mean(e < 0.0) > acc.mean()
    lacc, lseg, counts = pm.device.block_pockets(e, 0.0)
    assert np.array_equal(lacc, acc) and np.array_equal(lseg, seg) and counts[:2] == (n_seg, n_bl)
    rng = np.random.default_rng(1)
    g = pm.Grid(x.box, e.shape, acc, "accessibility", np.zeros(3))
    for xf in rng.random((2000, 3)):
        assert pm.accessible(g, xf) == pm.accessible(g, xf, (1, 1, 1)) == oacc.accessible(acc, xf)


def test_lta_replicated_kat(pm, orc, oacc, uff):
    x = pm.read_cif(golden("crystals", "LTA.cif"))
    acc1, _, _, _, _ = oacc.block(oracle_grid_kj(pm, orc, x, uff, pm.required_n_pts(x.box, 2.0)), 0.0)
    xr = pm.replicate(x, (2, 3, 1))
    er = oracle_grid_kj(pm, orc, xr, uff, pm.required_n_pts(xr.box, 2.0))
    accr, _, _, n_bl, _ = oacc.block(er, 0.0)
    assert acc1.shape == (7, 7, 7) and np.array_equal(acc1, accr[:7, :7, :7])
    assert n_bl > 0
    rng = np.random.default_rng(7)
    g1 = pm.Grid(x.box, acc1.shape, acc1, "accessibility", np.zeros(3))
    gr = pm.Grid(xr.box, accr.shape, accr, "accessibility", np.zeros(3))
    for xf in rng.random((10000, 3)):
        assert pm.accessible(gr, xf) == pm.accessible(g1, xf, (2, 3, 1))


@pytest.mark.parametrize("name", ["SBMOF-1", "CAXVII_clean"])
def test_no_pockets_kat(pm, orc, oacc, uff, name):
    x = pm.read_cif(golden("crystals", f"{name}.cif"))
    e = oracle_grid_kj(pm, orc, x, uff, pm.required_n_pts(x.box, 1.0))
    acc, _, _, n_bl, _ = oacc.block(e, 5.0)
    assert n_bl == 0
    assert np.array_equal(acc, e < 5.0)
    lacc, _, counts = pm.device.block_pockets(e, 5.0)
    assert counts[1] == 0 and np.array_equal(lacc, acc)
