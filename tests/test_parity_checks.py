"""The parity checks of tests/_parity.py catch the mistakes they are there
for.  Mistakes are planted in copies of the oracle's own energies and
derivatives (on the CPU): the checks must pass the oracle's values unchanged
(and moved by rounding-sized amounts) and fail each planted error, in
particular an error on the insertion that carries a block's K_H that the
condition-scale bound alone would let through."""
import math

import numpy as np
import pytest

from _parity import (TOL, block_sums_from_energies, check_block_sums, check_energies, derivative_sum_bounds,
                     repulsive_block_allowance, weight_ratio, weighted_mask)
from conftest import golden


@pytest.fixture(scope="module")
def co2_stream(pm, orc):
    """CO2 in CAXVII_clean (Dreiding, r_c 12.5 A, Ewald 1e-6): 5 x 300
    insertions with attractive ones in every block."""
    x = pm.read_cif(golden("crystals", "CAXVII_clean.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    co2 = pm.Molecule.from_dir(golden("molecules", "CO2"))
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), lj_off.T, [ff.index(s) for s in co2.atom_species], q_off.T,
                         co2.charges_q, ewald_precision=1e-6)
    nb = 300
    v, e = o.widom_energies(2024, 0, 5 * nb)
    s = o.widom_scales(2024, 0, 5 * nb)
    return nb, v, e, s


def test_unchanged_energies_pass(co2_stream):
    nb, v, e, s = co2_stream
    groups = np.repeat(np.arange(5), nb)
    rd, rw, rs = check_energies(v, e, v, e, TOL["mixed"], "same", cond_scale=s, groups=groups)
    assert rd == 0.0 and rw == 0.0 and rs == 0.0
    want = block_sums_from_energies(v, e, nb, 298.0)
    assert check_block_sums(want, want, 1e-9, "same") == (0.0, 0.0)


def test_weighted_error_caught_where_condition_scale_passes(co2_stream):
    """An error of 3e-5 (|E| + T) on each block's lowest (attractive)
    insertion passes the condition-scale bound (S is many times |E| for
    attractive insertions) and fails the weighted bound."""
    nb, v, e, s = co2_stream
    groups = np.repeat(np.arange(5), nb)
    tot = v + e
    bad_e = e.copy()
    lows = []
    for b in range(5):
        blk = np.arange(b * nb, (b + 1) * nb)
        i = blk[np.argmin(np.where(np.isfinite(tot[blk]), tot[blk], np.inf))]
        if tot[i] < 0:
            lows.append(i)
            bad_e[i] += 3e-5 * (abs(v[i]) + abs(e[i]) + 298.0)
    assert len(lows) == 5
    # the condition-scale bound alone accepts it
    fin = np.isfinite(tot)
    err_s = np.abs((v + bad_e)[fin] - tot[fin]) / (s[fin] + 1.0)
    assert err_s.max() < TOL["mixed"]
    with pytest.raises(AssertionError, match="weighted"):
        check_energies(v, bad_e, v, e, TOL["mixed"], "planted", cond_scale=s, groups=groups)


def test_weighted_bound_scales_with_the_boltzmann_weight(co2_stream):
    """An error of 3e-5 (|E| + T) on a block's second-lowest insertion fails
    when that insertion weighs more than a third of the lowest, and passes
    once it is e^-3 lighter: the bound follows the insertion's share of K_H."""
    nb, v, e, s = co2_stream
    tot = v + e
    groups = np.repeat(np.arange(5), nb)
    seen = set()
    for b in range(5):
        blk = np.arange(b * nb, (b + 1) * nb)
        fin = blk[np.isfinite(tot[blk])]
        order = fin[np.argsort(tot[fin])]
        i0, i1 = order[0], order[1]
        gap = (tot[i1] - tot[i0]) / 298.0
        bad_e = e.copy()
        bad_e[i1] += 3e-5 * (abs(v[i1]) + abs(e[i1]) + 298.0)
        if gap < 1.0:
            with pytest.raises(AssertionError, match="weighted"):
                check_energies(v, bad_e, v, e, TOL["mixed"], "planted", groups=groups)
            seen.add("heavy")
        elif gap > 3.0:
            check_energies(v, bad_e, v, e, TOL["mixed"], "planted", groups=groups)
            seen.add("light")
    assert seen, "no block with a clear second-lowest insertion"


def test_repulsive_error_caught_by_condition_scale(co2_stream):
    nb, v, e, s = co2_stream
    tot = v + e
    fin = np.isfinite(tot)
    i = int(np.flatnonzero(fin)[np.argmax(tot[fin])])
    assert tot[i] > 1e4
    bad_v = v.copy()
    bad_v[i] += 2e-5 * (s[i] + 1.0)
    with pytest.raises(AssertionError, match="condition scale"):
        check_energies(bad_v, e, v, e, TOL["mixed"], "planted", cond_scale=s, groups=np.repeat(np.arange(5), nb))


def test_overlap_flags_and_block_sums(co2_stream):
    nb, v, e, s = co2_stream
    tot = v + e
    fin = np.flatnonzero(np.isfinite(tot))
    inf = np.flatnonzero(~np.isfinite(tot))
    assert len(inf) > 0
    bad = v.copy()
    bad[inf[0]] = 1e9
    with pytest.raises(AssertionError, match="overlap"):
        check_energies(bad, e, v, e, TOL["mixed"], "planted")
    bad = v.copy()
    bad[fin[0]] = np.inf
    with pytest.raises(AssertionError, match="overlap"):
        check_energies(bad, e, v, e, TOL["mixed"], "planted")
    want = block_sums_from_energies(v, e, nb, 298.0)
    got = want.copy()
    got[2, 0] *= 1.0 + 2e-4
    with pytest.raises(AssertionError, match="K_H"):
        check_block_sums(got, want, 1e-4, "planted")
    got = want.copy()
    got[3, 1] += 2e-4 * abs(want[3, 1] + want[3, 2])
    with pytest.raises(AssertionError, match="<U>"):
        check_block_sums(got, want, 1e-4, "planted")


def test_block_sums_follow_the_reference_rules():
    """src/henry.jl:236-243: Inf energies weigh 0 and stay out of W."""
    v = np.array([-100.0, np.inf, 50.0, 10.0])
    e = np.array([-20.0, 5.0, np.inf, -5.0])
    got = block_sums_from_energies(v, e, 4, 300.0)[0]
    w0, w3 = math.exp(120.0 / 300.0), math.exp(-5.0 / 300.0)
    assert np.allclose(got, [w0 + w3, -100.0 * w0 + 10.0 * w3, -20.0 * w0 - 5.0 * w3], rtol=1e-15)


def test_weight_ratio():
    tot = np.array([-100.0, -100.0 + 298.0, np.inf, 5.0, 5.0 + 2 * 298.0])
    groups = np.array([0, 0, 0, 1, 1])
    r, lows = weight_ratio(tot, groups)
    assert lows.tolist() == [0, 3]
    assert np.allclose(r, [1.0, math.exp(-1.0), 0.0, 1.0, math.exp(-2.0)], rtol=1e-15)


def test_weighted_mask_and_repulsive_allowance():
    tot = np.array([-5000.0, -5000.0 + 40 * 298.0 - 1.0, -5000.0 + 40 * 298.0 + 1.0, np.inf,
                    3e4, 5e4, 1e9, np.inf])
    groups = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    m = weighted_mask(tot, groups)
    assert m.tolist() == [True, True, False, False, True, False, False, False]
    allow, u_allow = repulsive_block_allowance(tot, np.zeros_like(tot), 4, 1e-5, 298.0)
    assert allow[0] == 0.0 and 1e-5 < u_allow[0] < 1e-4
    assert allow[1] == pytest.approx(math.expm1(1e-5 * (3e4 + 298.0) / 298.0), rel=1e-6)
    assert u_allow[1] == pytest.approx(1e-5 * (3e4 + 298.0) / 3e4, rel=1e-6)
    # a <U> that cancels from large pair terms gets the allowance its terms imply
    _, u_c = repulsive_block_allowance(np.array([4330.0]), np.array([-4279.0]), 1, 1e-5, 298.0)
    assert u_c[0] == pytest.approx(1e-5 * (4330.0 + 4279.0 + 298.0) / 51.0, rel=1e-9)


@pytest.fixture(scope="module")
def co2_dual(pm, orc):
    """The CO2 workload's framework (5 charged species), 5 x 2000 insertions:
    the dual oracle's derivatives and their summand conditions."""
    case = pm.synthetic.co2_widom_case()
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), lj_off.T, [ff.index(s) for s in co2.atom_species], q_off.T,
                         co2.charges_q, ewald_precision=case.ewald_precision)
    pe = np.array([ff.pure_epsilon[s] for s in ff.species])
    ps = np.array([ff.pure_sigma[s] for s in ff.species])
    ft = sorted({ff.index(s) for s in x.species})
    nb = 2000
    v, e, d, cd = o.widom_param_grad(31, 0, 5 * nb, pe, ps, ft, cond=True)
    v2, e2, d2 = o.widom_param_grad(31, 0, 5 * nb, pe, ps, ft)
    assert np.array_equal(d, d2) and np.array_equal(v, v2) and np.array_equal(e, e2)
    return nb, v, e, d, cd


def _sums(nb, v, e, d):
    tot = v + e
    fin = np.isfinite(tot)
    w = np.where(fin, np.exp(-np.where(fin, tot, 0.0) / 298.0), 0.0)
    G = (w[:, None, None] * d).reshape(-1, nb, *d.shape[1:]).sum(axis=1)
    return G, -G.sum(axis=0) / (298.0 * w.sum())


def test_derivative_condition_bounds_the_derivative(co2_dual):
    """A sum of summands is no larger than the sum of their magnitudes, and a
    neutral probe's charge derivatives cancel across its sites: their
    condition is far above the derivative itself."""
    nb, v, e, d, cd = co2_dual
    fin = np.isfinite(v + e)
    assert np.all(cd[fin] >= np.abs(d[fin]) * (1.0 - 1e-12))
    assert np.all(cd[~fin] == 0.0)
    G, _ = _sums(nb, v, e, d)
    tot = v + e
    w = np.where(fin, np.exp(-np.where(fin, tot, 0.0) / 298.0), 0.0)
    C = (w[:, None, None] * cd).sum(axis=0)
    assert np.max(C[:, 2] / np.abs(G.sum(axis=0)[:, 2])) > 100.0   # q_t: cancellation across sites
    assert np.max(C[:, 0] / np.abs(G.sum(axis=0)[:, 0])) < 10.0    # eps_t: none to speak of


def test_derivative_bounds_pass_rounding_sized_errors(co2_dual):
    """Every summand of every derivative moved by up to 0.5 TOL of itself (as
    cond bounds) and every energy by 0.5 TOL (|E_vdw| + |E_es| + T): within
    the bounds."""
    nb, v, e, d, cd = co2_dual
    bb, bg = derivative_sum_bounds(v, e, d, cd, nb, TOL["mixed"], 298.0)
    rng = np.random.Generator(np.random.PCG64(5))
    fin = np.isfinite(v + e)
    dd = d + 0.5 * TOL["mixed"] * cd * rng.choice([-1.0, 1.0], size=d.shape)
    ee = e + np.where(fin, 0.5 * TOL["mixed"] * (np.abs(np.where(fin, v, 0.0)) + np.abs(e) + 298.0), 0.0)
    G0, g0 = _sums(nb, v, e, d)
    G1, g1 = _sums(nb, v, ee, dd)
    assert np.all(np.abs(G1 - G0) <= bb)
    assert np.all(np.abs(g1 - g0) <= bg)


@pytest.mark.parametrize("mistake", ["eps sign", "sigma 1%", "q 1%"])
def test_derivative_bounds_catch_planted_errors(co2_dual, mistake):
    """A sign error in one species' d/d eps, a 1% error in every d/d sigma,
    a 1% error in every d/d q: each fails the bounds."""
    nb, v, e, d, cd = co2_dual
    bb, bg = derivative_sum_bounds(v, e, d, cd, nb, TOL["mixed"], 298.0)
    bad = d.copy()
    if mistake == "eps sign":
        bad[:, 1, 0] = -bad[:, 1, 0]
    elif mistake == "sigma 1%":
        bad[:, :, 1] *= 1.01
    else:
        bad[:, :, 2] *= 1.01
    G0, g0 = _sums(nb, v, e, d)
    G1, g1 = _sums(nb, v, e, bad)
    assert np.max(np.abs(G1 - G0) / bb) > 1.0
    assert np.max(np.abs(g1 - g0) / bg) > 1.0
