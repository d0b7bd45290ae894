"""Large energy grids at full size: one 3000-atom triclinic framework on a
200^3 grid with a unit point-charge probe and a methane probe, and 80^3 grids
of smaller frameworks.  The whole grid runs on the device; the oracle checks
4096 sampled grid points of each plus the 64 lowest-energy points of the
device grid (its cost grows with the sample, not the grid), to the bounds of
tests/_parity.py (the sample's lowest point, every point weighted by its
Boltzmann factor at 298 K, every point against its condition scale; 1e-5
mixed, 1e-10 FP64)."""
import math

import numpy as np
import pytest

from _parity import TOL, check_energies, describe, say

pytestmark = pytest.mark.gpu


def _oracle(orc, pm, case, probe):
    x, ff = case.crystal, case.ljff
    lj_off, q_off = pm.site_offsets(probe)
    return orc.OracleSystem(
        (x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
        [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table, ff.sigma2_table,
        math.sqrt(ff.r2_cutoff), lj_off.T, [ff.index(s) for s in probe.atom_species], q_off.T,
        probe.charges_q, ewald_precision=case.ewald_precision)


def _sampled_check(pm, o, grid_data, n_pts, probe, precision, seed, n_sample=4096, n_low=64):
    rng = np.random.Generator(np.random.PCG64(seed))
    n1, n2, n3 = n_pts
    idx = rng.integers(0, [n1, n2, n3], size=(n_sample, 3))
    flat = np.where(np.isfinite(grid_data), grid_data, np.inf).ravel(order="F")
    low = np.argpartition(flat, n_low)[:n_low]
    idx = np.concatenate([idx, np.stack(np.unravel_index(low, n_pts, order="F"), axis=1)])
    rep = np.array(o.rep, dtype=float)
    m = len(idx)
    want_v, want_e, scale, got = np.zeros(m), np.zeros(m), np.zeros(m), np.zeros(m)
    for p, (i, j, k) in enumerate(idx):
        x = np.array([i / (n1 - 1), j / (n2 - 1), k / (n3 - 1)]) / rep
        if probe.n_atoms:
            want_v[p] = o.vdw(x.reshape(1, 3))
            scale[p] += o.vdw_scale(x.reshape(1, 3))
        if probe.n_charges:
            sr, lr = o.phi(x)
            q = probe.charges_q[0]
            want_e[p] = q * (sr + lr)
            scale[p] += abs(q) * o.phi_scale(x)
        got[p] = grid_data[i, j, k]
    # the device grid holds totals; the oracle's split sets the |E_vdw| + |E_es| scale
    return check_energies(got, 0.0, want_v, want_e, TOL[precision], "grid points", cond_scale=scale)


@pytest.fixture(scope="module")
def large(pm, orc):
    case = pm.synthetic.large_grid_case()
    return case


@pytest.mark.parametrize("probe_name", ["ion", "CH4"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_large_grid_200(pm, orc, large, probe_name, precision):
    n_pts = large.meta["n_pts"]
    say(f"case: 200^3 grid, 3000-atom framework, probe {probe_name} ({precision})")
    probe = large.probes[probe_name]
    g = pm.energy_grid(large.crystal, probe, large.ljff, resolution=n_pts, units="K", verbose=False,
                       precision=precision, electrostatics=True)
    ms, _ = pm.device.last_kernel_stats()
    pk, cand = pm.device.last_pair_kernel_stats()
    assert g.data.shape == n_pts
    o = _oracle(orc, pm, large, probe)
    res = _sampled_check(pm, o, g.data, n_pts, probe, precision, seed=7)
    npts = float(np.prod(n_pts))
    say(f"   {npts / (ms * 1e-3):.3e} grid points/s (device {ms:.2f} ms, pair kernel {pk:.2f} ms, "
        f"{cand / npts:.0f} candidates/point); {describe(res)}")
    if precision == "fp64":
        # the faces at fractional 0 and 1 coincide (test/grid.jl:43-55)
        d = g.data.copy()
        d[~np.isfinite(d) | (np.abs(d) > 1e6)] = 0.0
        np.testing.assert_allclose(d[0], d[-1], atol=1e-7, rtol=1e-9)
        np.testing.assert_allclose(d[:, :, 0], d[:, :, -1], atol=1e-7, rtol=1e-9)


def test_small_grids_80(pm, orc):
    """80^3 grids of 32 frameworks (200..1500 atoms): all on the device, four
    of them checked at sampled points."""
    cases = pm.synthetic.small_grid_cases()
    total_pts, total_ms, worst = 0.0, 0.0, 0.0
    for c in cases:
        say(f"case: 80^3 grid, framework {c.meta['index']} ({c.crystal.n_atoms} atoms), ion probe")
        probe = c.probes["ion"]
        g = pm.energy_grid(c.crystal, probe, c.ljff, resolution=c.meta["n_pts"], units="K",
                           verbose=False, precision="mixed", electrostatics=True)
        total_ms += pm.device.last_kernel_stats()[0]
        total_pts += float(np.prod(c.meta["n_pts"]))
        if c.meta["index"] % 8 == 0:
            o = _oracle(orc, pm, c, probe)
            _, rw, _ = _sampled_check(pm, o, g.data, c.meta["n_pts"], probe, "mixed", seed=c.meta["index"],
                                      n_sample=512)
            worst = max(worst, rw)
    say(f"   32 x 80^3 point-charge grids: {total_pts / (total_ms * 1e-3):.3e} grid points/s; "
        f"max weighted error {worst:.2e} of |E|+T")


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_grids_into_caller_buffers(pm, precision):
    """pm_energy_grid and pm_energy_grid_batch writing into buffers the caller
    reuses give the same grids, bit for bit, as into new ones."""
    from porousmaterials_jl_amd.grid import _grid_setup
    cases = pm.synthetic.small_grid_cases()[:4]
    say(f"case: 4 80^3 ion grids into reused host buffers ({precision})")
    runs = [_grid_setup(c.crystal, c.probes["ion"], c.ljff, c.meta["n_pts"], "K", False, False,
                        c.ewald_precision, electrostatics=True)[0] for c in cases]
    want = [pm.device.energy_grid(*r, precision) for r in runs]
    bufs = [np.full_like(w, np.nan) for w in want]
    for _ in range(2):
        got = pm.device.energy_grid_batch(runs, precision, outs=bufs)
        assert all(np.shares_memory(g, b) for g, b in zip(got, bufs))
        assert all(np.array_equal(b, w) for b, w in zip(bufs, want))
    buf = np.full_like(want[0], np.nan)
    assert np.shares_memory(pm.device.energy_grid(*runs[0], precision, out=buf), buf)
    assert np.array_equal(buf, want[0])
