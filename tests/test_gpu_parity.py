"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (stated per test, checks in tests/_parity.py), TOL = 1e-10
(PM_FP64) or 1e-5 (PM_MIXED), per group of energies (a Widom block, a set of
poses, a grid):
* every energy, weighted by its Boltzmann weight relative to the group's
  largest at T = 298 K: (w/w_max) |E_gpu - E_oracle| <= TOL * (|E_vdw| +
  |E_es| + T), which bounds its effect on K_H (the lowest energy of a group
  to TOL (|E| + T));
* every finite energy: |E_gpu - E_oracle| <= TOL * (S + 1 K), where S is the
  oracle's condition scale of the energy, the sum of the absolute pair
  contributions behind it, sum |LJ_ij| + sum_s |q_s| (sum_j |erfc term_sj| +
  |Fourier_s|).  A sum of terms each computed to relative error e is accurate
  to e * S, not to e * |E| (a repulsive E can cancel from terms of 10^6 K);
* Widom K_H and <U> per block: 1e-4 relative (mixed), 1e-9 (FP64) against
  the oracle's, except blocks whose every insertion is repulsive (held to
  the change of S that their energies' tolerance implies);
* overlap (+Inf) flags identical.
"""
import math

import numpy as np
import pytest

from _parity import (TOL, block_sums_from_energies, check_block_sums, check_energies, describe,
                     repulsive_block_allowance, say)
from conftest import golden

pytestmark = pytest.mark.gpu

KH_RTOL = {"fp64": 1e-9, "mixed": 1e-4}


def _lattice(box):
    return (box.a, box.b, box.c, box.alpha, box.beta, box.gamma)


def _oracle(orc, pm, xtal, ff, molecule, rep=None, ewald_sr_cutoff=None, eps=1e-6):
    lj_off, q_off = pm.site_offsets(molecule)
    return orc.OracleSystem(
        _lattice(xtal.box), xtal.xf.T, [ff.index(s) for s in xtal.species], xtal.atom_q,
        ff.epsilon_table, ff.sigma2_table, math.sqrt(ff.r2_cutoff), lj_off.T,
        [ff.index(s) for s in molecule.atom_species], q_off.T, molecule.charges_q,
        ewald_precision=eps, rep=rep, ewald_sr_cutoff=ewald_sr_cutoff)


# ------------------------------------------------------------- KATs ----
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_xe_sbmof1_kat(pm, orc, precision):
    say(f"case: Xe/SBMOF-1 Dreiding r_c 12.5 ({precision})")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    xr = pm.replicate(x, pm.replication_factors(x.box, ff))
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    o = _oracle(orc, pm, x, ff, xe)
    for cart, kat in (([0.0, 0.0, 0.0], -5041.58), ([0.494265, 2.22668, 0.450354], 12945.838)):
        m = pm.Frac(xe, xr.box)
        pm.translate_to(m, xr.box.c_to_f @ np.array(cart))
        e = pm.vdw_energy(xr, m, ff, precision=precision)
        want = o.vdw(m.atoms_x.T)
        check_energies(e, 0.0, want, 0.0, TOL[precision], f"Xe at {cart}", cond_scale=o.vdw_scale(m.atoms_x.T))
        assert abs(e - kat) < 0.005 + 1e-5 * abs(kat)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_doctest_full_precision(pm, precision):
    """docs/src/force_field.md:185-194 printed 5.73882798944654e6."""
    say(f"case: Xe/SBMOF-1 UFF unreplicated doctest ({precision})")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    xe = pm.Frac(pm.Molecule.from_dir(golden("molecules", "Xe")), x.box)
    pm.translate_to(xe, x.box.c_to_f @ np.array([0.0, 1.0, 3.0]))
    e = pm.vdw_energy(x, xe, ff, precision=precision)
    assert math.isclose(e, 5.73882798944654e6, rel_tol=1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nu1000_ion_kat(pm, orc, precision):
    """test/electrostatics.jl:33-46 through the device path."""
    say(f"case: ion q=0.8096 in NU-1000 ({precision})")
    nu = pm.read_cif(golden("crystals", "NU-1000_Greg.cif"))
    rep = pm.replication_factors(nu.box, 12.5)
    xr = pm.replicate(nu, rep)
    ep = pm.setup_Ewald_sum(xr.box, 12.5, eps=1e-6)
    species = sorted(set(nu.species))
    ff = pm.LJForceField("none", {s: 3.0 for s in species}, {s: 0.0 for s in species}, 12.5)
    o = _oracle(orc, pm, nu, ff, pm.point_charge(0.8096), rep=rep)
    for cart, kat, atol in (([9.535619863743, 20.685576379935, 0.127344239990], 111373.38, 2.5),
                            ([4.269654927228, 23.137319129548, 28.352847101096], -531.0, 0.5),
                            ([-0.047382031804, 7.209555961450, 5.158180463556], -2676.8230141, 0.5)):
        xf = xr.box.c_to_f @ np.array(cart)
        es = pm.electrostatic_potential_energy(xr, pm.ion(0.8096, xf), ep, precision=precision)
        sr, lr = o.phi(xf)
        check_energies(0.0, es.sr + es.lr, 0.0, 0.8096 * (sr + lr), TOL[precision], f"ion at {cart}",
                       cond_scale=0.8096 * o.phi_scale(xf))
        assert abs(pm.total(es) - kat) < atol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zif71_co2_kat(pm, precision):
    """test/electrostatics.jl:130-153 through the device path."""
    say(f"case: CO2-EPM2 in ZIF-71 ({precision})")
    z = pm.read_cif(golden("crystals", "zif71_bogus_charges.cif"))
    pm.strip_numbers_from_atom_labels(z)
    ff = pm.LJForceField.from_csv(golden("forcefields", "Greg_bogus_ZIF71.csv"), r_cutoff=12.8)
    co2 = pm.Frac(pm.Molecule.from_dir(golden("molecules", "CO2EPM2")), z.box)
    pose = np.array([[0.50543, 0.57349, 0.50788], [0.46884, 0.57393, 0.52461],
                     [0.54203, 0.57305, 0.49116]]).T
    co2.atoms_x = pose.copy()
    co2.charges_x = pose.copy()
    assert abs(pm.vdw_energy(z, co2, ff, precision=precision) - (-132.56)) < 0.01
    ep = pm.setup_Ewald_sum(z.box, 12.0, eps=1e-6)
    es = pm.electrostatic_potential_energy(z, co2, ep, precision=precision)
    assert abs(pm.total(es) - (-9.37846564)) < 0.1
    tot = pm.total_electrostatic_potential_energy(z, [co2, co2], ep, precision=precision)
    assert math.isclose(pm.total(tot), 2 * pm.total(es), rel_tol=1e-12)


# --------------------------------------------------- synthetic CO2 ----
@pytest.fixture(scope="module")
def co2case(pm, orc):
    case = pm.synthetic.co2_widom_case()
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision)
    o = _oracle(orc, pm, x, ff, co2)
    return case, xr, ep, o


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_poses(pm, orc, co2case, precision):
    """2048 CO2 poses from the insertion stream: vdw and Ewald per pose."""
    say(f"case: 2048 CO2 poses, synthetic 600-atom triclinic, ({precision})")
    case, xr, ep, o = co2case
    co2 = case.probes["CO2"]
    lj_off, q_off = pm.site_offsets(co2)
    n = 2048
    poses = np.zeros((n, 3, 3))
    for i in range(n):
        u = orc.insertion_uniforms(777, i)
        R = orc.arvo_rotation(u[3:])
        poses[i] = u[:3] + (o.c_to_f @ (R @ lj_off)).T
    vdw = pm.vdw_energies(xr, co2.atom_species, poses, case.ljff, precision=precision)
    sr, lr = pm.electrostatic_energies(xr, co2.charges_q, poses, ep, precision=precision)
    want_v = np.array([o.vdw(p) for p in poses])
    phis = [[o.phi(p[s]) for s in range(3)] for p in poses]
    want_sr = np.array([sum(q * ph[s][0] for s, q in enumerate(co2.charges_q)) for ph in phis])
    want_lr = np.array([sum(q * ph[s][1] for s, q in enumerate(co2.charges_q)) for ph in phis])
    fin = np.isfinite(want_v)
    assert 0 < fin.sum() < n
    s_v = np.array([o.vdw_scale(p) for p in poses])
    s_e = np.array([sum(abs(q) * o.phi_scale(p[s]) for s, q in enumerate(co2.charges_q)) for p in poses])
    res = check_energies(vdw, sr + lr, want_v, want_sr + want_lr, TOL[precision], "poses", cond_scale=s_v + s_e)
    _, _, ev = check_energies(vdw, 0.0, want_v, 0.0, TOL[precision], "vdw", cond_scale=s_v)
    _, _, es = check_energies(0.0, sr + lr, 0.0, want_sr + want_lr, TOL[precision], "Ewald", cond_scale=s_e)
    fl = np.abs(lr - want_lr) / (np.abs(want_lr) + 1.0)
    assert fl.max() <= TOL[precision], f"Fourier: {fl.max():.3e}"
    say(f"   {describe(res)} (vdw {ev:.2e}, Ewald {es:.2e} of S+1 K); Fourier {fl.max():.2e} of |E_lr|+1 K")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_widom_insertions_match_oracle(pm, co2case, precision):
    """Per-insertion energies of pm_widom and its block sums vs the oracle
    replaying the same Philox stream (5 blocks x 400 insertions)."""
    say(f"case: Widom CO2 5x400 insertions vs oracle ({precision})")
    case, xr, ep, o = co2case
    ff = case.ljff
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(case.probes["CO2"], ff)
    nb = 400
    seed = 2024
    sums, en = pm.device.widom(fw, ff, ew, pr, 298.0, seed, 5, nb, precision, return_energies=True)
    v, e = o.widom_energies(seed, 0, 5 * nb)
    scale = o.widom_scales(seed, 0, 5 * nb)
    res = check_energies(en[:, 0], en[:, 1], v, e, TOL[precision], "Widom E", cond_scale=scale,
                         groups=np.repeat(np.arange(5), nb))
    want = block_sums_from_energies(v, e, nb, 298.0)
    allow, u_allow = repulsive_block_allowance(v, e, nb, TOL[precision], 298.0)
    es, eu = check_block_sums(sums, want, KH_RTOL[precision], f"Widom blocks ({precision})", allow, u_allow)
    say(f"   {describe(res)}; blocks: K_H {es:.2e}, <U> {eu:.2e}; all-repulsive blocks {int(np.sum(allow > 0))}")


@pytest.mark.parametrize("eps", [1e-8, 1e-10])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_widom_energies_tight_ewald(pm, orc, precision, eps):
    """Per-insertion energies at tighter Ewald precisions (alpha r_c 3.7 and
    4.3 against 3.1 at 1e-6, so each segment of the erfc tables spans more of
    erfc): 5 x 120 CO2 insertions vs the oracle at that precision."""
    say(f"case: Widom CO2 5x120 at Ewald precision {eps:g} vs oracle ({precision})")
    case = pm.synthetic.co2_widom_case()
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    xr = pm.replicate(x, pm.replication_factors(x.box, ff))
    ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=eps)
    o = _oracle(orc, pm, x, ff, co2, eps=eps)
    assert math.isclose(o.alpha, ep.alpha, rel_tol=1e-12)
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(co2, ff)
    nb, seed = 120, 99
    _, en = pm.device.widom(fw, ff, ew, pr, 298.0, seed, 5, nb, precision, return_energies=True)
    v, e = o.widom_energies(seed, 0, 5 * nb)
    scale = o.widom_scales(seed, 0, 5 * nb)
    res = check_energies(en[:, 0], en[:, 1], v, e, TOL[precision], "Widom E", cond_scale=scale,
                         groups=np.repeat(np.arange(5), nb))
    say(f"   alpha r_c {ep.alpha * math.sqrt(ff.r2_cutoff):.2f}: {describe(res)}")


def test_widom_deterministic(pm, co2case):
    say("case: Widom determinism (two identical mixed runs)")
    case, xr, ep, o = co2case
    ff = case.ljff
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(case.probes["CO2"], ff)
    a = pm.device.widom(fw, ff, ew, pr, 298.0, 1, 5, 3000, "mixed")
    b = pm.device.widom(fw, ff, ew, pr, 298.0, 1, 5, 3000, "mixed")
    assert np.array_equal(a, b)


def test_widom_full_size_mixed_vs_fp64(pm, co2case):
    """Full workload (10^6 insertions): K_H and <U> of the mixed path within
    1e-4 of the FP64 path on the same stream."""
    say("case: Widom CO2 5 x 200000, mixed vs FP64")
    case, xr, ep, o = co2case
    ff = case.ljff
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(case.probes["CO2"], ff)
    nb = case.meta["n_per_block"]
    seed = case.meta["insertion_seed"]
    s64 = pm.device.widom(fw, ff, ew, pr, 298.0, seed, 5, nb, "fp64")
    s32 = pm.device.widom(fw, ff, ew, pr, 298.0, seed, 5, nb, "mixed")
    r64 = pm.block_statistics(s64, nb, 298.0, 1000.0)
    r32 = pm.block_statistics(s32, nb, 298.0, 1000.0)
    for k in ("henry coefficient [mol/(m³-bar)]", "⟨U⟩ (K)", "⟨U, vdw⟩ (K)", "⟨U, es⟩ (K)"):
        assert math.isclose(r32[k], r64[k], rel_tol=1e-4), (k, r32[k], r64[k])
    np.testing.assert_allclose(s32[:, 0], s64[:, 0], rtol=1e-4)


def test_henry_coefficient_api(pm, orc, tmp_path):
    """henry_coefficient end to end on CO2-EPM2 in ZIF-71 (test/henry.jl
    style): the results dict from the device block sums equals the
    statistics of the oracle's energies on the same insertion stream."""
    say("case: henry_coefficient API, CO2-EPM2 in ZIF-71")
    z = pm.read_cif(golden("crystals", "zif71_bogus_charges.cif"))
    pm.strip_numbers_from_atom_labels(z)
    ff = pm.LJForceField.from_csv(golden("forcefields", "Greg_bogus_ZIF71.csv"), r_cutoff=12.8)
    co2 = pm.Molecule.from_dir(golden("molecules", "CO2EPM2"))
    pm.henry.PATHS["simulations"] = str(tmp_path)
    ipv = 3000 / z.box.Omega
    r, sums = pm.henry_coefficient(z, co2, 298.0, ff, insertions_per_volume=ipv, verbose=False,
                                   autosave=True, seed=31, precision="fp64", return_block_sums=True)
    assert len(r) == 16
    nb = math.ceil(max(5, math.ceil(ipv * z.box.Omega)) / 5)
    o = _oracle(orc, pm, z, ff, co2)
    v, e = o.widom_energies(31, 0, 5 * nb)
    want = np.array([o.block_sums(v[b * nb:(b + 1) * nb], e[b * nb:(b + 1) * nb], 298.0) for b in range(5)])
    rw = pm.block_statistics(want, nb, 298.0, pm.crystal_density(z))
    for k in ("henry coefficient [mol/(kg-Pa)]", "⟨U⟩ (kJ/mol)", "Qst (kJ/mol)", "err ⟨U, es⟩ (kJ/mol)",
              "⟨U, es⟩ (kJ/mol)", "err henry coefficient [mmol/(g-bar)]"):
        assert math.isclose(r[k], rw[k], rel_tol=1e-9), (k, r[k], rw[k])
    assert r["⟨U, es⟩ (kJ/mol)"] != 0.0 and np.isfinite(r["⟨U⟩ (kJ/mol)"])
    assert any(p.suffix == ".json" for p in tmp_path.iterdir())
    # mixed precision through the same API within 1e-4
    r32 = pm.henry_coefficient(z, co2, 298.0, ff, insertions_per_volume=ipv, verbose=False,
                               autosave=False, seed=31, precision="mixed")
    for k in ("henry coefficient [mol/(kg-Pa)]", "⟨U⟩ (kJ/mol)"):
        assert math.isclose(r32[k], r[k], rel_tol=1e-4), (k, r32[k], r[k])


# ------------------------------------------------------------ grids ----
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_grid_full_size(pm, orc, precision):
    """Methane grid, 424-atom cubic framework, 50^3 points (the full
    workload) against the oracle."""
    say(f"case: CH4 grid 50^3, 424 atoms ({precision})")
    case = pm.synthetic.lj_grid_case()
    x, ff = case.crystal, case.ljff
    ch4 = case.probes["CH4"]
    g = pm.energy_grid(x, ch4, ff, resolution=case.meta["n_pts"], units="K", verbose=False,
                       precision=precision)
    o = _oracle(orc, pm, x, ff, ch4)
    want = o.energy_grid(case.meta["n_pts"]).reshape(case.meta["n_pts"], order="F")
    scale = o.energy_grid_scales(case.meta["n_pts"])
    res = check_energies(g.data.ravel(order="F"), 0.0, want.ravel(order="F"), 0.0, TOL[precision], "grid",
                         cond_scale=scale)
    say(f"   {describe(res)}")
    gk = pm.energy_grid(x, ch4, ff, resolution=case.meta["n_pts"], units="kJ_mol", verbose=False,
                        precision=precision)
    fin = np.isfinite(want)
    np.testing.assert_allclose(gk.data[fin], g.data[fin] * 8.314 / 1000.0, rtol=1e-15)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_point_charge_grid(pm, orc, precision):
    """Unit point-charge grid in a charged triclinic framework (per-point
    Ewald energy) against the oracle, 12^3 points."""
    say(f"case: point-charge grid 12^3, 600 atoms ({precision})")
    case = pm.synthetic.co2_widom_case(n_per_block=1)
    x, ff = case.crystal, case.ljff
    probe = pm.point_charge(1.0)
    g = pm.energy_grid(x, probe, ff, resolution=(12, 12, 12), units="K", verbose=False,
                       precision=precision, electrostatics=True)
    o = _oracle(orc, pm, x, ff, probe)
    want = o.energy_grid((12, 12, 12))
    scale = o.energy_grid_scales((12, 12, 12))
    res = check_energies(0.0, g.data.ravel(order="F"), 0.0, want, TOL[precision], "charge grid", cond_scale=scale)
    say(f"   {describe(res)}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_charge_grid_last_erfc_segment(pm, orc, precision):
    """Ion-probe grid points whose distance to a charge lies in the last
    segment of the erfc tables, [r_c - h, r_c) with h = r_c / 512, against the
    oracle.  A +1/-1 pair in a 40 A cubic box (no replication) and 3131 points
    along a, spaced 40/3130 A, put points at 12.78 and 12.79 A from the +1
    charge.  Reading the last segment at its start instead of at the point's
    own offset is an error of about 1e-7 of the condition scale there."""
    say(f"case: 3131 x 2 x 2 ion grid through the last erfc segment ({precision})")
    ff = pm.LJForceField("e", {"A": 3.0, "B": 3.0}, {"A": 50.0, "B": 50.0}, 12.8)
    box = pm.Box(40.0, 40.0, 40.0, math.pi / 2, math.pi / 2, math.pi / 2)
    x = pm.Crystal("pair", box, ["A", "B"], np.array([[0.0, 0.5], [0.0, 0.5], [0.0, 0.5]]), atom_q=[1.0, -1.0])
    probe = pm.point_charge(1.0)
    n = (3131, 2, 2)
    g = pm.energy_grid(x, probe, ff, resolution=n, units="K", verbose=False, precision=precision,
                       electrostatics=True)
    o = _oracle(orc, pm, x, ff, probe)
    assert tuple(o.rep) == (1, 1, 1)
    r = np.arange(n[0]) * 40.0 / (n[0] - 1)
    last = (r >= 12.8 * 511 / 512) & (r < 12.8)
    assert last.sum() == 2
    want = o.energy_grid(n)
    scale = o.energy_grid_scales(n)
    res = check_energies(0.0, g.data.ravel(order="F"), 0.0, want, TOL[precision], "last-segment grid",
                         cond_scale=scale)
    say(f"   {describe(res)}")


# ------------------------------------------------------- edge cases ----
def test_edge_cases(pm):
    say("case: edge cases (empty framework, uncharged framework, zero poses, bad species)")
    ff = pm.LJForceField("e", {"A": 3.0, "B": 3.4}, {"A": 50.0, "B": 30.0}, 12.8)
    box = pm.Box(26.0, 26.0, 26.0, math.pi / 2, math.pi / 2, math.pi / 2)
    empty = pm.Crystal("empty", box, [], np.zeros((3, 0)))
    m = pm.Frac(pm.lj_sphere("A"), box)
    assert pm.vdw_energy(empty, m, ff) == 0.0
    un = pm.Crystal("one", box, ["A"], np.array([[0.5], [0.5], [0.5]]))
    ep = pm.setup_Ewald_sum(box, 12.8)
    es = pm.electrostatic_potential_energy(un, pm.ion(1.0, [0.1, 0.1, 0.1]), ep)
    assert es.sr == 0.0 and es.lr == 0.0
    # exact overlap -> Inf in both precisions
    on = pm.Frac(pm.lj_sphere("B"), box)
    pm.translate_to(on, [0.5, 0.5, 0.5])
    assert math.isinf(pm.vdw_energy(un, on, ff, precision="fp64"))
    assert math.isinf(pm.vdw_energy(un, on, ff, precision="mixed"))
    # zero poses
    assert pm.vdw_energies(un, ["A"], np.zeros((0, 1, 3)), ff).size == 0
    # species outside the force field
    ff2 = pm.LJForceField("small", {"A": 3.0}, {"A": 50.0}, 12.8)
    with pytest.raises(KeyError):
        pm.vdw_energy(un, on, ff2)
    from porousmaterials_jl_amd import _lib
    fw = pm.device.framework(un, ff)
    small = pm.device.forcefield(ff2)
    with pytest.raises(_lib.PMError, match="outside the force field"):
        pm.device.vdw_energies(fw, small, [5], np.zeros((1, 1, 3)), "fp64")
    with pytest.raises(ValueError):
        pm.device.precision_code("fp16")
