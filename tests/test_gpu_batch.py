"""Batched screening (pm_widom_batch) and batched grids (pm_energy_grid_batch).

A run's energies are a function of the run alone (pair vectors from absolute
Cartesian coordinates, FP32 partial sums ordered by the framework's image
list), so a batch must reproduce every run's stand-alone result bit for bit,
whatever the batch composition, scheduling order or launch-group split.  The
screening regime itself (small cells replicated up to 5x5x5, thousands of
k-vectors) is checked against the oracle per insertion, to the tolerances of
tests/_parity.py."""
import math

import numpy as np
import pytest

from _parity import TOL, check_energies, describe, say

pytestmark = pytest.mark.gpu


def _handles(pm, case, probe_name):
    x, ff = case.crystal, case.ljff
    probe = case.probes[probe_name]
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    fw = pm.device.framework(xr, ff)
    ew = None
    if probe.n_charges:
        ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision, verbose=False)
        ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(probe, ff)
    return fw, pm.device.forcefield(ff), ew, pr, xr


@pytest.fixture(scope="module")
def screen(pm):
    return pm.synthetic.screening_cases(n_frameworks=24)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_widom_batch_equals_single_runs(pm, screen, precision):
    say(f"case: pm_widom_batch of 24 frameworks x (CH4, CO2), 5 x 1000 insertions ({precision})")
    runs = []
    for i, c in enumerate(screen):
        for name in ("CH4", "CO2"):
            fw, ff, ew, pr, _ = _handles(pm, c, name)
            runs.append((fw, ff, ew, pr, 298.0, 1000 + i, 5, 1000))
    got = pm.device.widom_batch(runs, precision)
    for run, g in zip(runs, got):
        want = pm.device.widom(*run, precision)
        assert np.array_equal(g, want), (run[5], g, want)


def test_widom_batch_split_across_launch_groups(pm):
    """Eight CO2 runs of 5 x 150,000 insertions (1.8e7 point slots) exceed one
    launch group (2^24 slots): one run is split between two groups."""
    say("case: pm_widom_batch 8 x CO2 5 x 150000 (two launch groups), mixed")
    case = pm.synthetic.co2_widom_case()
    fw, ff, ew, pr, _ = _handles(pm, case, "CO2")
    runs = [(fw, ff, ew, pr, 298.0, 77 + k, 5, 150000) for k in range(8)]
    got = pm.device.widom_batch(runs, "mixed")
    for k in (0, 3, 5, 7):
        want = pm.device.widom(*runs[k], "mixed")
        assert np.array_equal(got[k], want), k


def test_screening_regime_vs_oracle(pm, orc, screen):
    """Per-insertion CO2 energies in the smallest and the largest replicated
    screening frameworks vs the oracle, to the bounds of tests/_parity.py
    (FP64 1e-10, mixed 1e-5)."""
    sizes = [(pm.replication_factors(c.crystal.box, c.ljff), c) for c in screen]
    by_nrep = sorted(sizes, key=lambda t: int(np.prod(t[0])) * t[1].crystal.n_atoms)
    for rep, c in (by_nrep[0], by_nrep[-1]):
        x, ffl = c.crystal, c.ljff
        co2 = c.probes["CO2"]
        say(f"case: screening framework {c.meta['index']} ({x.n_atoms} atoms, rep {rep}), CO2 vs oracle")
        fw, ff, ew, pr, xr = _handles(pm, c, "CO2")
        lj_off, q_off = pm.site_offsets(co2)
        o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                             [ffl.index(s) for s in x.species], x.atom_q, ffl.epsilon_table,
                             ffl.sigma2_table, math.sqrt(ffl.r2_cutoff), lj_off.T,
                             [ffl.index(s) for s in co2.atom_species], q_off.T, co2.charges_q,
                             ewald_precision=c.ewald_precision)
        assert tuple(o.rep) == tuple(rep)
        nb = 60
        v, e = o.widom_energies(31, 0, 5 * nb)
        scale = o.widom_scales(31, 0, 5 * nb)
        for precision in ("fp64", "mixed"):
            _, en = pm.device.widom(fw, ff, ew, pr, 298.0, 31, 5, nb, precision, return_energies=True)
            res = check_energies(en[:, 0], en[:, 1], v, e, TOL[precision], f"screening ({precision})",
                                 cond_scale=scale, groups=np.repeat(np.arange(5), nb))
            say(f"   {precision}: {describe(res)}")


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_grid_batch_equals_single_grids(pm, precision):
    say(f"case: pm_energy_grid_batch of 6 frameworks x (ion, CH4) at 24^3 ({precision})")
    cases = pm.synthetic.small_grid_cases(n_frameworks=6, n_pts=24)
    runs = []
    for c in cases:
        for name in ("ion", "CH4"):
            fw, ff, ew, pr, _ = _handles(pm, c, name)
            runs.append((fw, ff, ew, pr, c.meta["n_pts"]))
    got = pm.device.energy_grid_batch(runs, precision)
    for run, g in zip(runs, got):
        want = pm.device.energy_grid(*run, precision)
        assert np.array_equal(g, want)


def test_henry_coefficients_api(pm):
    """The screening call surface: henry_coefficients over (crystal, molecule)
    pairs equals henry_coefficient per pair."""
    say("case: henry_coefficients over 4 (crystal, molecule) pairs")
    cases = pm.synthetic.screening_cases(n_frameworks=2)
    pairs = [(c.crystal, c.probes[n]) for c in cases for n in ("CH4", "CO2")]
    ff = cases[0].ljff
    res = pm.henry_coefficients(pairs, 298.0, ff, insertions_per_volume=2.0, seeds=[5, 6, 7, 8])
    for (x, m), seed, r in zip(pairs, [5, 6, 7, 8], res):
        want = pm.henry_coefficient(x, m, 298.0, ff, insertions_per_volume=2.0, seed=seed, verbose=False,
                                    autosave=False)
        k = "henry coefficient [mol/(kg-Pa)]"
        assert r[k] == want[k]
        assert r["⟨U⟩ (K)"] == want["⟨U⟩ (K)"] or (math.isnan(r["⟨U⟩ (K)"]) and math.isnan(want["⟨U⟩ (K)"]))


def test_batch_rejects_bad_runs(pm):
    say("case: pm_widom_batch argument errors")
    from porousmaterials_jl_amd._lib import PMError
    case = pm.synthetic.co2_widom_case()
    fw, ff, ew, pr, _ = _handles(pm, case, "CO2")
    with pytest.raises(PMError, match="run 1"):
        pm.device.widom_batch([(fw, ff, ew, pr, 298.0, 1, 5, 10), (fw, ff, ew, pr, -1.0, 1, 5, 10)], "mixed")
    with pytest.raises(PMError, match="Ewald"):
        pm.device.widom_batch([(fw, ff, None, pr, 298.0, 1, 5, 10)], "mixed")


def test_chunk_tuning_changes_no_result(pm):
    """Per-point energies do not depend on the cell layout: pm_widom gives the
    same block sums bit for bit under any (cell points, chunk radius) the
    image-list path serves, and after pm_autotune_widom."""
    say("case: CO2 5 x 20000 under six chunk tunings and after the autotuner (mixed)")
    case = pm.synthetic.co2_widom_case()
    fw, ff, ew, pr, _ = _handles(pm, case, "CO2")
    before = pm.device.get_tuning()
    try:
        pm.device.set_tuning(640, 3.0)
        want = pm.device.widom(fw, ff, ew, pr, 298.0, 11, 5, 20000, "mixed")
        # (4096: 4 x 4 columns of about 19,000 points, two stacked cells of 146 slabs each)
        for cp, cr in ((192, 3.0), (256, 3.0), (1280, 3.0), (1280, 6.0), (384, 4.5), (4096, 6.0)):
            pm.device.set_tuning(cp, cr)
            got = pm.device.widom(fw, ff, ew, pr, 298.0, 11, 5, 20000, "mixed")
            assert np.array_equal(got, want), (cp, cr)
        best = pm.device.autotune_widom(fw, ff, ew, pr, 298.0, 11, 5, 20000, "mixed")
        say(f"   autotuner: {best}")
        assert np.array_equal(pm.device.widom(fw, ff, ew, pr, 298.0, 11, 5, 20000, "mixed"), want)
    finally:
        pm.device.set_tuning(before["cell_points"], before["chunk_r_max"])
