"""Adjoint Widom pass (pm_widom_gradient) on the CO2 workload's framework
(CO2, and probes of one and of five sites).

d ln K_H / d theta for theta = (eps_t, sigma_t, q_t) of the 5 framework
species is checked against central finite differences (h = 1e-4 relative) of
the oracle's FP64 ln K_H on the same insertion stream, to 1e-4 relative
(floor: 1e-6 of the largest component).  q_t is a uniform shift of every
atom of species t (h = 1e-4 of their mean |q|).  The per-block weighted
derivative sums sum_i w_i dE_i/d theta are checked against the oracle's
dual-number derivatives (oracle/dual.c, the forward-mode check) on the same
stream, on the image-list and the all-atoms pair paths: FP64 to 1e-8
relative (floor 1e-8 of the block's largest component); mixed within the
derivative-sum bounds of tests/_parity.py (TOL 1e-5 per summand and per
energy).  The pass's block sums must equal pm_widom's bit for bit."""
import math
import sys

import numpy as np
import pytest

from _parity import TINY_S, TOL, derivative_sum_bounds

pytestmark = pytest.mark.gpu

T = 298.0
N_BLOCKS, NB = 5, 400
SEED = 2024


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def adj(pm):
    return _handles(pm, pm.synthetic.co2_widom_case())


def _oracle_lnS(orc, pm, case, ljff, q_atoms):
    x, co2 = case.crystal, case.probes["CO2"]
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ljff.index(s) for s in x.species], q_atoms, ljff.epsilon_table, ljff.sigma2_table,
                         math.sqrt(ljff.r2_cutoff), lj_off.T, [ljff.index(s) for s in co2.atom_species], q_off.T,
                         co2.charges_q, ewald_precision=case.ewald_precision)
    v, e = o.widom_energies(SEED, 0, N_BLOCKS * NB)
    E = v + e
    fin = np.isfinite(E)
    return math.log(np.sum(np.exp(-E[fin] / T)))


def _perturbed_ff(pm, ff, species, which, factor):
    eps = dict(ff.pure_epsilon)
    sig = dict(ff.pure_sigma)
    if which == "epsilon":
        eps[species] *= factor
    else:
        sig[species] *= factor
    return pm.LJForceField(ff.name, sig, eps, math.sqrt(ff.r2_cutoff), ff.mixing_rules)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_pass_block_sums_equal_pm_widom(pm, adj, precision):
    say(f"case: pm_widom_gradient block sums vs pm_widom, CO2 5 x 3000 ({precision})")
    case, fw, ew, pr = adj
    s_g, _ = pm.device.widom_gradient(fw, case.ljff, ew, pr, T, 7, 5, 3000, precision, case.ljff)
    s_w = pm.device.widom(fw, case.ljff, ew, pr, T, 7, 5, 3000, precision)
    assert np.array_equal(s_g, s_w)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dlnKH_matches_oracle_finite_differences(pm, orc, adj, precision):
    say(f"case: d ln K_H / d(eps_t, sigma_t, q_t), CO2 5 x {NB} insertions, vs oracle central FD ({precision})")
    case, fw, ew, pr = adj
    ff, x = case.ljff, case.crystal
    sums, grads = pm.device.widom_gradient(fw, ff, ew, pr, T, SEED, N_BLOCKS, NB, precision, ff)
    S = sums[:, 0].sum()
    assert S > 0.0
    g_dev = -grads.sum(axis=0) / (T * S)  # (n_species, 3)
    species = sorted(set(x.species), key=ff.index)
    h = 1e-4
    q0 = np.array(x.atom_q, dtype=float)
    fd, dev = [], []
    for sp in species:
        k = ff.index(sp)
        for j, which in enumerate(("epsilon", "sigma")):
            p0 = ff.pure_epsilon[sp] if which == "epsilon" else ff.pure_sigma[sp]
            lp = _oracle_lnS(orc, pm, case, _perturbed_ff(pm, ff, sp, which, 1 + h), q0)
            lm = _oracle_lnS(orc, pm, case, _perturbed_ff(pm, ff, sp, which, 1 - h), q0)
            fd.append((lp - lm) / (2 * h * p0))
            dev.append(g_dev[k, j])
        mask = np.array([s == sp for s in x.species])
        hq = h * float(np.mean(np.abs(q0[mask])))
        lp = _oracle_lnS(orc, pm, case, ff, q0 + hq * mask)
        lm = _oracle_lnS(orc, pm, case, ff, q0 - hq * mask)
        fd.append((lp - lm) / (2 * hq))
        dev.append(g_dev[k, 2])
    fd, dev = np.array(fd), np.array(dev)
    floor = 1e-6 * np.max(np.abs(fd))
    err = np.abs(dev - fd) / np.maximum(np.abs(fd), floor)
    for sp_i, sp in enumerate(species):
        say(f"   {sp}: d/d eps {dev[3 * sp_i]:+.6e} (FD {fd[3 * sp_i]:+.6e}), d/d sigma {dev[3 * sp_i + 1]:+.6e} "
            f"(FD {fd[3 * sp_i + 1]:+.6e}), d/d q {dev[3 * sp_i + 2]:+.6e} (FD {fd[3 * sp_i + 2]:+.6e})")
    say(f"   max relative deviation from FD: {err.max():.2e}")
    assert err.max() <= 1e-4


def _handles(pm, case, probe="CO2"):
    x, ff = case.crystal, case.ljff
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision, verbose=False)
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(case.probes[probe], ff)
    return case, fw, ew, pr


def _dual_check(pm, orc, adj, precision, nb, seed, probe="CO2"):
    case, fw, ew, pr = adj
    ff, x, co2 = case.ljff, case.crystal, case.probes[probe]
    sums, grads = pm.device.widom_gradient(fw, ff, ew, pr, T, seed, N_BLOCKS, nb, precision, ff)
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), lj_off.T, [ff.index(s) for s in co2.atom_species], q_off.T,
                         co2.charges_q, ewald_precision=case.ewald_precision)
    pe = np.array([ff.pure_epsilon[s] for s in ff.species])
    ps = np.array([ff.pure_sigma[s] for s in ff.species])
    ft = sorted({ff.index(s) for s in x.species})
    v, e, d, cd = o.widom_param_grad(seed, 0, N_BLOCKS * nb, pe, ps, ft, cond=True)
    E = v + e
    fin = np.isfinite(E)
    w = np.where(fin, np.exp(-np.where(fin, E, 0.0) / T), 0.0)
    want = np.stack([(w[b * nb:(b + 1) * nb, None, None] * d[b * nb:(b + 1) * nb]).sum(axis=0)
                     for b in range(N_BLOCKS)])                      # (blocks, species, 3)
    got = grads[:, ft, :]
    assert np.all(grads[:, [k for k in range(grads.shape[1]) if k not in ft], :] == 0.0)
    S = sums[:, 0].sum()
    g_dev = -got.sum(axis=0) / (T * S)
    g_dual = -want.sum(axis=0) / (T * w.sum())
    dl_floor = 1e-8 if precision == "fp64" else 1e-6
    dl = np.abs(g_dev - g_dual) / np.maximum(np.abs(g_dual), dl_floor * np.abs(g_dual).max())
    if precision == "fp64":
        worst = 0.0
        for b in range(N_BLOCKS):
            scale = np.maximum(np.abs(want[b]), 1e-8 * np.abs(want[b]).max())
            worst = max(worst, float((np.abs(got[b] - want[b]) / scale).max()))
        say(f"   per-block weighted derivative sums: max relative deviation {worst:.2e}; d ln K_H/d theta: "
            f"{dl.max():.2e}")
        assert worst <= 1e-8 and dl.max() <= 1e-8
    else:
        bb, bg = derivative_sum_bounds(v, e, d, cd, nb, TOL["mixed"], T)
        # (a block whose weights all underflow has zero sums and a zero bound)
        rb = float((np.abs(got - want) / np.maximum(bb, TINY_S)).max())
        # (a component with a zero bound, e.g. the charge derivatives of an
        # uncharged probe, must match exactly)
        rg = float((np.abs(g_dev - g_dual) / np.maximum(bg, np.finfo(float).tiny)).max())
        say(f"   d ln K_H/d theta max relative deviation {dl.max():.2e}; error / bound: per-block sums {rb:.2e}, "
            f"d ln K_H/d theta {rg:.2e}")
        assert rb <= 1.0 and rg <= 1.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_block_derivative_sums_match_dual_oracle(pm, orc, adj, precision):
    say(f"case: sum_i w_i dE_i/d theta per block, CO2 5 x 2000 insertions, vs the dual-number oracle ({precision})")
    _dual_check(pm, orc, adj, precision, 2000, 31)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_all_atoms_path_derivative_sums_match_dual_oracle(pm, orc, adj, precision):
    """The pair kernel's all-atoms path (one cell per run: the chunk radius
    exceeds the image-list path's limit), where a candidate's type changes
    from atom to atom."""
    say(f"case: all-atoms pair path, CO2 5 x 2000 insertions, vs the dual-number oracle ({precision})")
    before = pm.device.get_tuning()
    try:
        pm.device.set_tuning(65536, 0.5)
        _dual_check(pm, orc, adj, precision, 2000, 37)
    finally:
        pm.device.set_tuning(before["cell_points"], before["chunk_r_max"])


@pytest.mark.parametrize("n_species", [1, 2, 3, 7, 8])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_kernels_for_every_species_count(pm, orc, precision, n_species):
    """Frameworks of 1, 2, 3, 7 and 8 species run the gradient pair kernels
    compiled for 2, 4 and 8 species and the Fourier kernels for 1, 2, 3, 7
    and 8: block sums equal pm_widom's bit for bit, derivative sums meet the
    dual-number oracle as above."""
    case = pm.synthetic.co2_widom_case(seed=40 + n_species, n_species=n_species)
    say(f"case: {n_species}-species framework, CO2 5 x 1000 insertions ({precision})")
    h = _handles(pm, case)
    _, fw, ew, pr = h
    s_g, _ = pm.device.widom_gradient(fw, case.ljff, ew, pr, T, 43, 5, 1000, precision, case.ljff)
    s_w = pm.device.widom(fw, case.ljff, ew, pr, T, 43, 5, 1000, precision)
    assert np.array_equal(s_g, s_w)
    _dual_check(pm, orc, h, precision, 1000, 43)


def _five_site_probe(pm):
    """CO2's three LJ atoms on the x axis and two off-atom charges +-0.4 e on
    the y axis: five sites, two of them charge-only."""
    ax = np.array([[0.0, -1.16, 1.16], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    qx = np.array([[0.0, 0.0], [0.55, -0.55], [0.0, 0.0]])
    return pm.Molecule("CO2_q5", ["C_CO2", "O_CO2", "O_CO2"], ax, [0.4, -0.4], qx, np.zeros(3))


@pytest.mark.parametrize("probe", ["CH4", "CO2_q5"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gradient_other_site_counts(pm, orc, precision, probe):
    """Probes of one site (CH4: uncharged, so no charge derivatives) and of
    five sites (three LJ-only, two charge-only) on the CO2 workload's
    framework: block sums equal pm_widom's bit for bit, derivative sums meet
    the dual-number oracle as above."""
    case = pm.synthetic.co2_widom_case()
    case.probes["CH4"] = pm.lj_sphere("CH4")
    case.probes["CO2_q5"] = _five_site_probe(pm)
    say(f"case: {probe} probe, 5 x 1000 insertions ({precision})")
    h = _handles(pm, case, probe)
    _, fw, ew, pr = h
    s_g, grads = pm.device.widom_gradient(fw, case.ljff, ew, pr, T, 47, 5, 1000, precision, case.ljff)
    s_w = pm.device.widom(fw, case.ljff, ew, pr, T, 47, 5, 1000, precision)
    assert np.array_equal(s_g, s_w)
    if probe == "CH4":
        assert np.all(grads[:, :, 2] == 0.0)
    _dual_check(pm, orc, h, precision, 1000, 47, probe)


def test_henry_coefficient_gradient_api(pm):
    """The host call surface on a real CIF: ZIF-71 (bogus charges) with
    CO2-EPM2 (test/electrostatics.jl:140-153 setting), mixed vs FP64 within
    1e-4 for every component larger than 1e-3 of the largest."""
    from conftest import golden
    say("case: henry_coefficient_gradient, ZIF-71 / CO2-EPM2, mixed vs FP64")
    x = pm.read_cif(golden("crystals", "zif71_bogus_charges.cif"))
    pm.strip_numbers_from_atom_labels(x)
    ff = pm.LJForceField.from_csv(golden("forcefields", "Greg_bogus_ZIF71.csv"), r_cutoff=12.8)
    m = pm.Molecule.from_dir(golden("molecules", "CO2EPM2"))
    r64, g64 = pm.henry_coefficient_gradient(x, m, 298.0, ff, insertions_per_volume=20, seed=3, precision="fp64")
    r32, g32 = pm.henry_coefficient_gradient(x, m, 298.0, ff, insertions_per_volume=20, seed=3, precision="mixed")
    want = pm.henry_coefficient(x, m, 298.0, ff, insertions_per_volume=20, seed=3, precision="fp64",
                                verbose=False, autosave=False)
    k = "henry coefficient [mol/(kg-Pa)]"
    assert r64[k] == want[k]
    a = np.array([[g64[s][p] for p in ("epsilon", "sigma", "charge")] for s in g64])
    b = np.array([[g32[s][p] for p in ("epsilon", "sigma", "charge")] for s in g32])
    big = np.abs(a) > 1e-3 * np.abs(a).max()
    say(f"   species {list(g64)}; max mixed-vs-FP64 relative deviation {np.max(np.abs(a - b)[big] / np.abs(a)[big]):.2e}")
    assert np.all(np.abs(a - b)[big] <= 1e-4 * np.abs(a)[big])
