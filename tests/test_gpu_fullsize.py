"""Widom runs at full size against the oracle's FP64 energies of the same
insertion stream (src/henry.jl:185-265 per insertion, :112-157 statistics).

* the synthetic CO2 workload: 600-atom triclinic cell replicated (2,2,2),
  rigid 3-site CO2 with Ewald (1e-6), 10^6 insertions = 5 x 200,000;
* Xe in SBMOF-1 (Dreiding, r_c 12.5 A, LJ only) and CO2 in CAXVII_clean
  (Dreiding, r_c 12.5 A, Ewald 1e-6), 5 x 200,000 insertions each
  (test/henry.jl:13-34), where K_H is a sum over many insertions rather than
  one;
* the screening set: all 256 frameworks x (CH4, CO2) in one pm_widom_batch
  call, 5 x 20 insertions per run;
* the adjoint at full size: d ln K_H / d(eps_t, sigma_t, q_t) of the CO2
  workload's 10^6 insertions against the oracle's dual-number derivatives.

Tolerances: per insertion as in tests/_parity.py (every insertion's
weight-scaled error to 1e-5 / 1e-10 of |E_vdw| + |E_es| + T, the first
20,000 insertions to the same fraction of their condition scale; the CO2
workload's dominant insertions to 1e-5 / 1e-10 of |E|); per block
K_H and <U> to 1e-4 (mixed) and 1e-9 (FP64) of the oracle's, with no
allowance derived from the energy tolerance; overlap flags identical.
"""
import math

import numpy as np
import pytest

from _parity import (TINY_S, TOL, block_sums_from_energies, check_block_sums, check_energies, derivative_sum_bounds,
                     describe, repulsive_block_allowance, say)
from conftest import golden

pytestmark = pytest.mark.gpu

KH_RTOL = {"fp64": 1e-9, "mixed": 1e-4}
N_BLOCKS, N_PER_BLOCK = 5, 200000


def _lattice(box):
    return (box.a, box.b, box.c, box.alpha, box.beta, box.gamma)


def _oracle(orc, pm, xtal, ff, molecule, eps=1e-6):
    lj_off, q_off = pm.site_offsets(molecule)
    return orc.OracleSystem(
        _lattice(xtal.box), xtal.xf.T, [ff.index(s) for s in xtal.species], xtal.atom_q,
        ff.epsilon_table, ff.sigma2_table, math.sqrt(ff.r2_cutoff), lj_off.T,
        [ff.index(s) for s in molecule.atom_species], q_off.T, molecule.charges_q, ewald_precision=eps)


def _device(pm, xtal, ff, molecule, eps=1e-6):
    xr = pm.replicate(xtal, pm.replication_factors(xtal.box, ff))
    fw = pm.device.framework(xr, ff)
    ew = None
    if molecule.n_charges and np.any(np.asarray(xtal.atom_q) != 0.0):
        ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=eps, verbose=False)
        ew = pm.device.ewald(xr, fw, ep)
    return fw, pm.device.forcefield(ff), ew, pm.device.probe(molecule, ff)


class FullRun:
    """Oracle energies of one full-size stream, computed once per module."""

    def __init__(self, pm, orc, name, xtal, ff, molecule, seed, eps=1e-6):
        self.name, self.xtal, self.ff, self.molecule, self.seed = name, xtal, ff, molecule, seed
        self.o = _oracle(orc, pm, xtal, ff, molecule, eps)
        say(f"case: {name}: oracle energies of {N_BLOCKS} x {N_PER_BLOCK} insertions "
            f"({orc.lib().orc_num_threads()} host threads)")
        self.v, self.e = self.o.widom_energies(seed, 0, N_BLOCKS * N_PER_BLOCK)
        self.n_scaled = 20000
        self.scale = self.o.widom_scales(seed, 0, self.n_scaled)
        self.sums = block_sums_from_energies(self.v, self.e, N_PER_BLOCK, 298.0)
        self.handles = _device(pm, xtal, ff, molecule, eps)

    def check(self, pm, precision):
        fw, ff, ew, pr = self.handles
        sums, en = pm.device.widom(fw, ff, ew, pr, 298.0, self.seed, N_BLOCKS, N_PER_BLOCK, precision,
                                   return_energies=True)
        groups = np.repeat(np.arange(N_BLOCKS), N_PER_BLOCK)
        res = check_energies(en[:, 0], en[:, 1], self.v, self.e, TOL[precision], f"{self.name} ({precision})",
                             groups=groups)
        n = self.n_scaled
        _, _, rel_s = check_energies(en[:n, 0], en[:n, 1], self.v[:n], self.e[:n], TOL[precision],
                                     f"{self.name} first {n} ({precision})", cond_scale=self.scale,
                                     groups=groups[:n])
        es, eu = check_block_sums(sums, self.sums, KH_RTOL[precision], f"{self.name} ({precision})")
        say(f"   {precision}: {describe(res)}; first {n}: {rel_s:.2e} of S+1 K; blocks: K_H {es:.2e}, "
            f"<U> {eu:.2e}")
        return sums, en


@pytest.fixture(scope="module")
def co2_full(pm, orc):
    case = pm.synthetic.co2_widom_case()
    return FullRun(pm, orc, "synthetic CO2, 600 atoms (2,2,2)", case.crystal, case.ljff, case.probes["CO2"],
                   case.meta["insertion_seed"], case.ewald_precision)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_co2_workload_full_size_vs_oracle(pm, co2_full, precision):
    """The CO2 workload at 10^6 insertions.  Its K_H is carried by one
    insertion per block (block 0: E = -13,711.6 K, vdw -1,415.2, es
    -12,296.4, at seed 1234567); that insertion is held to 1e-5 of |E|, i.e.
    0.14 K, on the mixed path and its error is printed."""
    r = co2_full
    sums, en = r.check(pm, precision)
    tot = r.v + r.e
    for b in range(N_BLOCKS):
        blk = slice(b * N_PER_BLOCK, (b + 1) * N_PER_BLOCK)
        i = b * N_PER_BLOCK + int(np.nanargmin(np.where(np.isfinite(tot[blk]), tot[blk], np.inf)))
        share = math.exp(-tot[i] / 298.0) / r.sums[b, 0]
        d = abs(en[i, 0] + en[i, 1] - tot[i])
        say(f"   block {b}: lowest insertion {i}: E {tot[i]:.4f} K (vdw {r.v[i]:.4f}, es {r.e[i]:.4f}), "
            f"{100 * share:.4f}% of the block's weight, device error {d:.2e} K")
        assert d <= TOL[precision] * abs(tot[i]), (b, i, d)
    if precision == "mixed":
        i0 = int(np.nanargmin(np.where(np.isfinite(tot[:N_PER_BLOCK]), tot[:N_PER_BLOCK], np.inf)))
        assert abs(tot[i0] - (-13711.6)) < 0.1, tot[i0]
        assert abs(en[i0, 0] + en[i0, 1] - tot[i0]) <= 0.14
    rs = pm.block_statistics(sums, N_PER_BLOCK, 298.0, 1000.0)
    ro = pm.block_statistics(r.sums, N_PER_BLOCK, 298.0, 1000.0)
    for k in ("henry coefficient [mol/(m³-bar)]", "⟨U⟩ (K)", "⟨U, vdw⟩ (K)", "⟨U, es⟩ (K)"):
        assert math.isclose(rs[k], ro[k], rel_tol=KH_RTOL[precision]), (k, rs[k], ro[k])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_co2_workload_full_size_cell_layouts(pm, co2_full, precision):
    """Every insertion's energies and every block sum of the 10^6-insertion
    CO2 workload are the same bit for bit under cube cells (64 points: one
    wave each), column cells (320: 20 x 23 columns of 102 slabs) and stacked
    column cells (4096: 9 x 10 x 3 cells of 174 slabs), all on the image-list
    path: the cell layout only schedules work."""
    fw, ff, ew, pr = co2_full.handles
    before = pm.device.get_tuning()
    try:
        got = []
        for cp in (64, 320, 4096):
            say(f"   {precision}: cell points {cp}")
            pm.device.set_tuning(cp, 6.0)
            got.append(pm.device.widom(fw, ff, ew, pr, 298.0, co2_full.seed, N_BLOCKS, N_PER_BLOCK, precision,
                                       return_energies=True))
        for cp, (sums, en) in zip((320, 4096), got[1:]):
            assert np.array_equal(sums, got[0][0]), cp
            assert np.array_equal(en, got[0][1]), cp
    finally:
        pm.device.set_tuning(before["cell_points"], before["chunk_r_max"])


@pytest.fixture(scope="module")
def co2_dual_full(pm, orc, co2_full):
    """The oracle's dual-number derivatives dE_i/d(eps_t, sigma_t, q_t) of
    every insertion of the CO2 workload's 10^6-insertion stream."""
    r = co2_full
    ff, x = r.ff, r.xtal
    pe = np.array([ff.pure_epsilon[s] for s in ff.species])
    ps = np.array([ff.pure_sigma[s] for s in ff.species])
    ft = sorted({ff.index(s) for s in x.species})
    say(f"case: CO2 workload: dual-number oracle derivatives of {N_BLOCKS} x {N_PER_BLOCK} insertions")
    v, e, d, cd = r.o.widom_param_grad(r.seed, 0, N_BLOCKS * N_PER_BLOCK, pe, ps, ft, cond=True)
    fin = np.isfinite(r.v)
    assert np.array_equal(np.isfinite(v), fin)
    assert np.max(np.abs(v[fin] + e[fin] - r.v[fin] - r.e[fin]) / (np.abs(r.v[fin]) + np.abs(r.e[fin]) + 1.0)) < 1e-10
    E = r.v + r.e
    w = np.where(fin, np.exp(-np.where(fin, E, 0.0) / 298.0), 0.0)
    G = np.stack([(w[b * N_PER_BLOCK:(b + 1) * N_PER_BLOCK, None, None]
                   * d[b * N_PER_BLOCK:(b + 1) * N_PER_BLOCK]).sum(axis=0) for b in range(N_BLOCKS)])
    bounds = derivative_sum_bounds(v, e, d, cd, N_PER_BLOCK, TOL["mixed"], 298.0)
    return ft, G, w.sum(), bounds


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_co2_workload_dlnKH_full_size_vs_dual_oracle(pm, co2_full, co2_dual_full, precision):
    """d ln K_H / d(eps_t, sigma_t, q_t) of the 5 framework species at the
    full 10^6 insertions (pm_widom_gradient) against the oracle's exact
    (dual-number) derivatives on the same stream.  FP64: 1e-8 relative, floor
    1e-6 of the largest component (per block: of the block's largest).
    Mixed: within the derivative-sum bounds of tests/_parity.py at TOL 1e-5
    (each summand of a derivative rounded to 1e-5 of itself, each weight moved
    by the energy tolerance): the charge derivatives of a neutral probe cancel
    across its sites by up to 10^4, which no FP32 pair arithmetic resolves to
    1e-4 of the result."""
    r = co2_full
    ft, G, S, (bb, bg) = co2_dual_full
    fw, ff, ew, pr = r.handles
    sums, grads = pm.device.widom_gradient(fw, ff, ew, pr, 298.0, r.seed, N_BLOCKS, N_PER_BLOCK, precision, r.ff)
    got = grads[:, ft, :]
    g_dev = -got.sum(axis=0) / (298.0 * sums[:, 0].sum())
    g_ref = -G.sum(axis=0) / (298.0 * S)
    err = np.abs(g_dev - g_ref) / np.maximum(np.abs(g_ref), 1e-6 * np.abs(g_ref).max())
    for i, t in enumerate(ft):
        say(f"   {r.ff.species[t]}: d ln K_H/d eps {g_dev[i, 0]:+.8e} (oracle {g_ref[i, 0]:+.8e}), d/d sigma "
            f"{g_dev[i, 1]:+.8e} ({g_ref[i, 1]:+.8e}), d/d q {g_dev[i, 2]:+.8e} ({g_ref[i, 2]:+.8e})")
    if precision == "fp64":
        blk = max(float((np.abs(got[b] - G[b]) / np.maximum(np.abs(G[b]), 1e-6 * np.abs(G[b]).max())).max())
                  for b in range(N_BLOCKS))
        say(f"   fp64: d ln K_H/d theta max relative deviation {err.max():.2e}; per-block sums {blk:.2e}")
        assert err.max() <= 1e-8 and blk <= 1e-8
    else:
        rb = float((np.abs(got - G) / np.maximum(bb, TINY_S)).max())
        rg = float((np.abs(g_dev - g_ref) / bg).max())
        say(f"   mixed: d ln K_H/d theta max relative deviation {err.max():.2e}; error / bound: per-block sums "
            f"{rb:.2e}, d ln K_H/d theta {rg:.2e}")
        assert rb <= 1.0 and rg <= 1.0


@pytest.fixture(scope="module")
def golden_full(pm, orc):
    out = {}
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    out["Xe"] = FullRun(pm, orc, "Xe in SBMOF-1 (Dreiding 12.5)", x, ff,
                        pm.Molecule.from_dir(golden("molecules", "Xe")), 4242)
    x = pm.read_cif(golden("crystals", "CAXVII_clean.cif"))
    out["CO2"] = FullRun(pm, orc, "CO2 in CAXVII_clean (Dreiding 12.5)", x, ff,
                         pm.Molecule.from_dir(golden("molecules", "CO2")), 4343)
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("which,kat", [("Xe", (0.00985348, -39.6811)), ("CO2", (2.88317e-5, -18.69582223))])
def test_golden_frameworks_full_size_vs_oracle(pm, golden_full, which, kat, precision):
    """Xe/SBMOF-1 and CO2/CAXVII_clean at 10^6 insertions vs the oracle, and
    the reference's statistical known answers (test/henry.jl:20-21,33-34,
    rtol 0.025) from the same run."""
    r = golden_full[which]
    sums, _ = r.check(pm, precision)
    rs = pm.block_statistics(sums, N_PER_BLOCK, 298.0, pm.crystal_density(r.xtal))
    ro = pm.block_statistics(r.sums, N_PER_BLOCK, 298.0, pm.crystal_density(r.xtal))
    for k in ("henry coefficient [mol/(kg-Pa)]", "⟨U⟩ (kJ/mol)"):
        assert math.isclose(rs[k], ro[k], rel_tol=KH_RTOL[precision]), (k, rs[k], ro[k])
    kh, u = rs["henry coefficient [mol/(kg-Pa)]"], rs["⟨U⟩ (kJ/mol)"]
    say(f"   K_H {kh:.6e} mol/(kg-Pa) (known answer {kat[0]:.6e}), <U> {u:.4f} kJ/mol ({kat[1]})")
    assert math.isclose(kh, kat[0], rel_tol=0.025)
    assert math.isclose(u, kat[1], rel_tol=0.025)


@pytest.fixture(scope="module")
def screening_oracle(pm, orc):
    cases = pm.synthetic.screening_cases()
    say(f"case: screening set: oracle energies of {len(cases)} frameworks x (CH4, CO2) x 5 x 20 insertions")
    runs = []
    for c in cases:
        for name in ("CH4", "CO2"):
            m = c.probes[name]
            o = _oracle(orc, pm, c.crystal, c.ljff, m, c.ewald_precision)
            seed = 50000 + 2 * c.meta["index"] + (name == "CO2")
            v, e = o.widom_energies(seed, 0, 100)
            runs.append((c, name, seed, v, e))
    return runs


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_screening_batch_every_run_vs_oracle(pm, screening_oracle, precision):
    """Every one of the 512 screening runs in one pm_widom_batch call: block
    K_H and <U> vs the oracle's; per-insertion energies of each run (the
    stand-alone run equals the batch bit for bit) vs the oracle.  These are
    dense frameworks and 20-insertion blocks: most blocks hold only repulsive
    insertions (E > 0, up to 10^7 K).  Blocks with an insertion at E <= 0 K
    are held to 1e-4 (mixed) in K_H; an all-repulsive block, whose K_H is
    e^(-E/T) of a pose with E ~ 10^5 K, to the change of S its energies'
    tolerance implies (the largest error over such blocks is printed).  <U> is
    held to 1e-4, or where a block's <U> cancels from much larger pair terms
    (E = 51 K from 4330 K of LJ and -4279 K of Ewald), to the change its
    energies' tolerance implies."""
    say(f"case: screening batch, 512 runs x 5 x 20 insertions vs oracle ({precision})")
    handles, runs = [], []
    for c, name, seed, v, e in screening_oracle:
        fw, ff, ew, pr = _device(pm, c.crystal, c.ljff, c.probes[name], c.ewald_precision)
        handles.append((fw, ff, ew, pr))
        runs.append((fw, ff, ew, pr, 298.0, seed, 5, 20))
    got = pm.device.widom_batch(runs, precision)
    worst = [0.0, 0.0, 0.0, 0.0, 0.0]
    n_zero = n_rep = 0
    for run, g, (c, name, seed, v, e) in zip(runs, got, screening_oracle):
        what = f"framework {c.meta['index']} {name}"
        _, en = pm.device.widom(*run, precision, return_energies=True)
        assert np.array_equal(pm.device.widom(*run, precision), g), what
        rd, rel_w, _ = check_energies(en[:, 0], en[:, 1], v, e, TOL[precision], what,
                                      groups=np.repeat(np.arange(5), 20))
        want = block_sums_from_energies(v, e, 20, 298.0)
        n_zero += int(np.sum(want[:, 0] == 0.0))
        allow, u_allow = repulsive_block_allowance(v, e, 20, TOL[precision], 298.0)
        rep = (allow > 0.0) & (want[:, 0] > 0.0)
        n_rep += int(rep.sum())
        attr = ~rep
        es, eu = check_block_sums(g[attr], want[attr], KH_RTOL[precision], what, None, u_allow[attr])
        es_r, eu_r = check_block_sums(g[rep], want[rep], KH_RTOL[precision], what + " (repulsive)", allow[rep],
                                      u_allow[rep])
        worst = [max(worst[0], rel_w), max(worst[1], es), max(worst[2], eu, eu_r), max(worst[3], es_r),
                 max(worst[4], rd)]
    say(f"   {precision}: dominant energies {worst[4]:.2e} of |E|+1 K, weighted {worst[0]:.2e} of |E|+T; "
        f"block K_H {worst[1]:.2e} "
        f"({n_rep} all-repulsive blocks: {worst[3]:.2e}), <U> {worst[2]:.2e}; "
        f"{n_zero} of 2560 blocks with every weight zero")
