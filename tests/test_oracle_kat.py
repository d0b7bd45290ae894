"""Pins the CPU oracle against the reference's known answers and against what
the mathematics fixes (closed forms, invariants, textbook constants).

Every expected number below is printed in the reference's tests or doctests
(file:line cited) or follows from mathematics stated in the test.
"""
import math
import os

import numpy as np
import pytest

from conftest import golden

PI = math.pi


def _lattice(box):
    return (box.a, box.b, box.c, box.alpha, box.beta, box.gamma)


def _sys(orc, pm, xtal, ff, lj_species, lj_off, q_off=None, q_site=None, rep=None,
         r_cutoff=None, ewald_sr_cutoff=None, ewald_precision=1e-6):
    return orc.OracleSystem(
        _lattice(xtal.box), xtal.xf.T, [ff.index(s) for s in xtal.species], xtal.atom_q,
        ff.epsilon_table, ff.sigma2_table,
        math.sqrt(ff.r2_cutoff) if r_cutoff is None else r_cutoff,
        np.zeros((0, 3)) if lj_off is None else lj_off,
        [ff.index(s) for s in lj_species],
        np.zeros((0, 3)) if q_off is None else q_off,
        np.zeros(0) if q_site is None else q_site,
        ewald_precision=ewald_precision, rep=rep, ewald_sr_cutoff=ewald_sr_cutoff)


# ----------------------------------------------------------- geometry ----
def test_box_convention_doctest(orc, pm):
    """docs/src/box.md / docs/src/crystal.md:48-54: Cartesian coordinates of
    SBMOF-1's first and last atoms as printed by the reference."""
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    F = orc.box_f_to_c(*_lattice(x.box))
    cart = F @ x.xf
    np.testing.assert_allclose(cart[:, 0], [4.594867082350715, 1.4395486200000005, 5.89964228469024],
                               rtol=1e-14)
    np.testing.assert_allclose(cart[:, -1], [-1.5321086078257065, 4.212266223, 16.239103154389543],
                               rtol=1e-14)
    # lower triangle exactly zero, a along x
    assert F[1, 0] == 0.0 and F[2, 0] == 0.0 and F[2, 1] == 0.0
    np.testing.assert_allclose(abs(np.linalg.det(F)), 1456.472102, rtol=1e-8)


def test_replication_factors_doctest(orc, pm):
    """docs/src/force_field.md:178-180: SBMOF-1 at r_c = 14 A -> (3, 6, 2)."""
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    assert orc.replication_factors(orc.box_f_to_c(*_lattice(x.box)), 14.0) == (3, 6, 2)


def test_replicate_order_and_box(orc):
    xf = np.array([[0.1, 0.2, 0.3], [0.9, 0.5, 0.05]])
    out = orc.replicate(xf, (2, 1, 3))
    assert out.shape == (12, 3)
    # ra outermost, then rb, rc, then atoms
    np.testing.assert_allclose(out[0], [0.05, 0.2, 0.1])
    np.testing.assert_allclose(out[1], [0.45, 0.5, 0.05 / 3])
    np.testing.assert_allclose(out[2], [0.05, 0.2, (0.3 + 1) / 3])
    np.testing.assert_allclose(out[6], [(0.1 + 1) / 2, 0.2, 0.1])


# ------------------------------------------------------ Lennard-Jones ----
def test_lennard_jones_closed_forms(orc):
    """test/vdw_energetics.jl:20-21: LJ(sigma) = 0, LJ(2^(1/6) sigma) = -eps."""
    s2, eps = 1.0, 3.0
    assert abs(orc.lennard_jones(s2, s2, eps)) < 1e-14
    assert math.isclose(orc.lennard_jones(2 ** (2 / 6) * s2, s2, eps), -eps, rel_tol=1e-14)
    # the 12-6 form itself at an arbitrary point
    r2, s2, e = 7.3, 9.61, 41.0
    assert math.isclose(orc.lennard_jones(r2, s2, e), 4 * e * ((s2 / r2) ** 6 - (s2 / r2) ** 3),
                        rel_tol=1e-13)


def _xe_sbmof1(pm):
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    return x, ff


def test_xe_sbmof1_raspa(orc, pm):
    """test/vdw_energetics.jl:26-41: Xe in replicated SBMOF-1, Dreiding,
    r_c 12.5: -5041.58 K at the origin, 12945.838 K at a Cartesian point
    (atol 0.005, matched to RASPA)."""
    x, ff = _xe_sbmof1(pm)
    s = _sys(orc, pm, x, ff, ["Xe"], np.zeros((1, 3)))
    e1 = s.vdw(np.zeros((1, 3)))
    assert abs(e1 - (-5041.58)) < 0.005
    xf = s.c_to_f @ np.array([0.494265, 2.22668, 0.450354])
    e2 = s.vdw(xf.reshape(1, 3))
    assert abs(e2 - 12945.838) < 0.005


def test_xe_sbmof1_doctest_full_precision(orc, pm):
    """docs/src/force_field.md:185-194: Xe at Cartesian (0, 1, 3) in the
    unreplicated SBMOF-1 with UFF (r_c 14) prints 5.73882798944654e6."""
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    s = _sys(orc, pm, x, ff, ["Xe"], np.zeros((1, 3)), rep=(1, 1, 1))
    xf = s.c_to_f @ np.array([0.0, 1.0, 3.0])
    e = s.vdw(xf.reshape(1, 3))
    assert math.isclose(e, 5.73882798944654e6, rel_tol=2e-15)


def test_overlap_is_inf(orc, pm):
    """src/vdw_energetics.jl:36-37: r^2 < 0.1 inside the cutoff -> Inf."""
    x, ff = _xe_sbmof1(pm)
    s = _sys(orc, pm, x, ff, ["Xe"], np.zeros((1, 3)))
    on_atom = s.xf[5] + np.array([1e-3, 0.0, 0.0]) / 23.0
    assert math.isinf(s.vdw(on_atom.reshape(1, 3)))


@pytest.mark.parametrize("c,expected", [(1, -4.3515e3), (2, -6.9000e2), (3, -1.1467e3), (4, -1.6790e1)])
def test_nist_lj_configurations(orc, c, expected):
    """test/vdw_energetics.jl:53-80: NIST LJ reference configurations,
    sigma = eps = 1, r_c = 3 sigma, atol 0.1."""
    with open(golden("nist", "lennardjones", f"lj_sample_config_periodic{c}.txt")) as f:
        lines = f.read().splitlines()
    dims = [float(v) for v in lines[0].split()]
    n = int(lines[1])
    xyz = np.array([[float(v) for v in lines[2 + i].split()[1:4]] for i in range(n)])
    F = orc.box_f_to_c(*dims, PI / 2, PI / 2, PI / 2)
    xf = xyz @ np.linalg.inv(F).T
    e = orc.total_vdw_gg(xf, F, 1.0, 1.0, 9.0)
    assert abs(e - expected) < 0.1


# ------------------------------------------------------------- Ewald ----
def test_ewald_params_nu1000(orc, pm):
    """test/electrostatics.jl:12-26: NU-1000, r_c 12.5, eps 1e-6 on the
    replicated box -> kreps (9, 9, 9), alpha ~ 0.2471."""
    x = pm.read_cif(golden("crystals", "NU-1000_Greg.cif"))
    F = orc.box_f_to_c(*_lattice(x.box))
    rep = orc.replication_factors(F, 12.5)
    b = x.box
    Fr = orc.box_f_to_c(b.a * rep[0], b.b * rep[1], b.c * rep[2], b.alpha, b.beta, b.gamma)
    alpha, k2 = orc.determine_ewald_params(12.5, 1e-6)
    assert orc.required_kreps(Fr, k2) == (9, 9, 9)
    assert abs(alpha - 0.2471) < 0.0001
    # alpha and max|k|^2 are the roots of the reference's error functions
    assert abs(math.erfc(12.5 * alpha) / 12.5 - 1e-6) < 1e-18
    assert abs(2 * math.exp(-k2 / (4 * alpha ** 2)) / k2 - 1e-6) < 1e-18


def test_required_kreps_cases(orc):
    """test/electrostatics.jl:27-31: reciprocal lattice (2, 10, 5)."""
    F = orc.box_f_to_c(0.5 * 2 * PI, 0.1 * 2 * PI, 0.2 * 2 * PI, PI / 2, PI / 2, PI / 2)
    assert orc.required_kreps(F, 2.1 ** 2) == (1, 0, 0)
    assert orc.required_kreps(F, 5.1 ** 2) == (2, 0, 1)
    assert orc.required_kreps(F, 10.1 ** 2) == (5, 1, 2)
    R = orc.reciprocal_lattice(F)
    np.testing.assert_allclose(R, np.diag([2.0, 10.0, 5.0]), atol=1e-13)


def test_kvector_enumeration(orc):
    """src/electrostatics.jl:204-229: half-space set, no (0,0,0), ka >= 0,
    ka = 0 -> kb >= 0, ka = kb = 0 -> kc > 0; together with -k it covers the
    full box exactly once."""
    F = orc.box_f_to_c(11.0, 13.0, 9.0, 1.4, 1.7, 1.9)
    kidx, wt = orc.precompute_kvec_wts((3, 2, 4), F, 0.3)
    full = {tuple(k) for k in kidx} | {tuple(-k) for k in kidx}
    assert len(full) == 2 * len(kidx)
    box = {(a, b, c) for a in range(-3, 4) for b in range(-2, 3) for c in range(-4, 5)} - {(0, 0, 0)}
    assert full == box
    # weight formula at one vector
    R = orc.reciprocal_lattice(F)
    k = R.T @ kidx[7]
    k2 = float(k @ k)
    omega = abs(np.linalg.det(F))
    assert math.isclose(wt[7], 2 * math.exp(-k2 / (4 * 0.09)) / k2 / orc.eps0() / omega, rel_tol=1e-14)


def test_unit_constant():
    """src/electrostatics.jl:6: 1/(4 pi eps0) in K A / e^2."""
    from oracle import oracle as orc
    assert math.isclose(1.0 / orc.four_pi_eps0(), 167100.95, rel_tol=1e-7)


def _nu1000(orc, pm, form):
    x = pm.read_cif(golden("crystals", "NU-1000_Greg.cif"))
    ff = pm.LJForceField("none", {"C": 3.0}, {"C": 0.0}, 12.5)
    sp = [0] * x.n_atoms
    s = orc.OracleSystem(_lattice(x.box), x.xf.T, sp, x.atom_q, np.zeros((1, 1)), np.ones((1, 1)),
                         12.5, np.zeros((0, 3)), [], np.zeros((1, 3)), [0.8096])
    return s


@pytest.mark.parametrize("cart,expected,atol", [
    ([9.535619863743, 20.685576379935, 0.127344239990], 111373.38, 2.5),
    ([4.269654927228, 23.137319129548, 28.352847101096], -531.0, 0.5),
    ([-0.047382031804, 7.209555961450, 5.158180463556], -2676.8230141, 0.5),
])
def test_nu1000_ion_energies(orc, pm, cart, expected, atol):
    """test/electrostatics.jl:33-46: a q = 0.8096 ion in NU-1000."""
    s = _nu1000(orc, pm, "eikr")
    xf = s.c_to_f @ np.array(cart)
    sr, lr = s.phi(xf, form="eikr")
    assert abs(0.8096 * (sr + lr) - expected) < atol
    sr2, lr2 = s.phi(xf, form="sk")
    # structure-factor form == reference form (same sum, reorganised)
    assert math.isclose(sr2, sr, rel_tol=1e-13)
    assert abs(lr2 - lr) < 1e-9 * max(1.0, abs(lr))


def test_zif71_co2_epm2(orc, pm):
    """test/electrostatics.jl:130-153: CO2-EPM2 at a fixed pose in ZIF-71
    with bogus charges: vdw -132.56 K (atol 0.01), Ewald -9.37846564 K
    (r_c 12.0, eps 1e-6; atol 0.1)."""
    x = pm.read_cif(golden("crystals", "zif71_bogus_charges.cif"))
    pm.strip_numbers_from_atom_labels(x)
    ff = pm.LJForceField.from_csv(golden("forcefields", "Greg_bogus_ZIF71.csv"), r_cutoff=12.8)
    co2 = pm.Molecule.from_dir(golden("molecules", "CO2EPM2"))
    assert co2.atom_species[0] == "C_CO2" and math.isclose(co2.charges_q[0], 0.7)
    pose = np.array([[0.50543, 0.57349, 0.50788], [0.46884, 0.57393, 0.52461],
                     [0.54203, 0.57305, 0.49116]])
    s = _sys(orc, pm, x, ff, co2.atom_species, np.zeros((3, 3)), np.zeros((3, 3)), co2.charges_q,
             rep=(1, 1, 1), ewald_sr_cutoff=12.0)
    assert abs(s.vdw(pose) - (-132.56)) < 0.01
    es_eikr = sum(q * sum(s.phi(p, form="eikr")) for q, p in zip(co2.charges_q, pose))
    es_sk = sum(q * sum(s.phi(p, form="sk")) for q, p in zip(co2.charges_q, pose))
    assert abs(es_eikr - (-9.37846564)) < 0.1
    assert abs(es_sk - es_eikr) < 1e-9


def _spce(orc, c):
    with open(golden("nist", "electrostatics", f"spce_sample_config_periodic{c}.txt")) as f:
        lines = f.read().splitlines()
    dims = [float(v) for v in lines[0].split()]
    n = int(lines[1]) * 3
    xyz, q, mol = [], [], []
    for i in range(n):
        t = lines[2 + i].split()
        xyz.append([float(v) for v in t[1:4]])
        q.append(-2 * 0.42380 if t[-1] == "O" else 0.42380)
        mol.append(i // 3)
    F = orc.box_f_to_c(*dims, PI / 2, PI / 2, PI / 2)
    return F, np.array(xyz) @ np.linalg.inv(F).T, np.array(q), np.array(mol), dims


@pytest.mark.parametrize("c,ref", [
    (1, {"real": -5.58889e05, "fourier": 6.27009e03, "self": -2.84469e06, "intra": 2.80999e06}),
    (2, {"real": -1.19295e06, "fourier": 6.03495e03, "self": -5.68938e06, "intra": 5.61998e06}),
])
def test_nist_spce_ewald_components(orc, c, ref):
    """test/electrostatics.jl:51-118: NIST SPC/E Ewald components with
    alpha = 5.6/L, kreps (5,5,5), |n|^2 < 27, r_c 10 (rtol 1e-5, intra 1e-4)."""
    F, xf, q, mol, dims = _spce(orc, c)
    alpha = 5.6 / dims[0]
    kidx, wt = orc.precompute_kvec_wts((5, 5, 5), F, alpha)
    keep = (kidx ** 2).sum(axis=1) < 27
    out = orc.ewald_gg_components(xf, q, mol, F, alpha, 10.0, kidx[keep], wt[keep])
    for k in ("real", "fourier", "self"):
        assert math.isclose(out[k], ref[k], rel_tol=1e-5), k
    assert math.isclose(out["intra"], ref["intra"], rel_tol=1e-4)


def test_madelung_rock_salt(orc):
    """Ewald real + Fourier + self energy of a rock-salt lattice reproduces
    the Madelung constant 1.747564594633 (textbook value)."""
    r0 = 2.82
    a = 2 * r0
    basis = []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                basis.append(((i, j, k), 1.0 if (i + j + k) % 2 == 0 else -1.0))
    pts, q = [], []
    n = 2  # 2x2x2 conventional cells -> 64 ions
    for cx in range(n):
        for cy in range(n):
            for cz in range(n):
                for (i, j, k), s in basis:
                    pts.append(((cx * 2 + i) / (2 * n), (cy * 2 + j) / (2 * n), (cz * 2 + k) / (2 * n)))
                    q.append(s)
    pts, q = np.array(pts), np.array(q)
    L = n * a
    F = orc.box_f_to_c(L, L, L, PI / 2, PI / 2, PI / 2)
    alpha = 3.8 / (L / 2)
    kidx, wt = orc.precompute_kvec_wts((12, 12, 12), F, alpha)
    out = orc.ewald_gg_components(pts, q, np.arange(len(q)), F, alpha, L / 2, kidx, wt)
    e = out["real"] + out["fourier"] + out["self"]
    madelung = -e * r0 * orc.four_pi_eps0() / (len(q) / 2)
    assert math.isclose(madelung, 1.747564594633, rel_tol=2e-6)


def test_ewald_alpha_invariance(orc):
    """Total Ewald energy of a neutral periodic system does not depend on the
    splitting parameter alpha (to the truncation error)."""
    rng = np.random.Generator(np.random.PCG64(5))
    n = 24
    xf = rng.random((n, 3))
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    F = orc.box_f_to_c(14.0, 15.0, 13.0, 1.45, 1.62, 1.70)
    widths = abs(np.linalg.det(F)) / np.linalg.norm(np.cross(F[:, [1, 2, 0]].T, F[:, [2, 0, 1]].T), axis=1)
    rc = 0.5 * widths.min()
    totals = []
    for alpha in (3.6 / rc, 4.0 / rc, 4.4 / rc):
        kmax2 = 4 * alpha * alpha * 30.0
        kreps = orc.required_kreps(F, kmax2)
        kidx, wt = orc.precompute_kvec_wts(kreps, F, alpha, kmax2)
        out = orc.ewald_gg_components(xf, q, np.arange(n), F, alpha, rc, kidx, wt)
        totals.append(out["real"] + out["fourier"] + out["self"])
    scale = np.sum(np.abs(q)) ** 2 / orc.four_pi_eps0() / rc
    assert max(totals) - min(totals) < 2e-5 * scale


def test_sk_form_equals_eikr_form_replicated(orc):
    """The unit-cell structure-factor Fourier sum equals the reference's
    e^{ik.r} sum over every replicated charge and replicated-box k-vector."""
    rng = np.random.Generator(np.random.PCG64(7))
    n = 40
    lat = (9.0, 10.5, 8.3, 1.50, 1.65, 1.72)
    xf = rng.random((n, 3))
    q = rng.uniform(-1.5, 1.5, n)
    q -= q.mean()
    s = orc.OracleSystem(lat, xf, [0] * n, q, np.zeros((1, 1)), np.ones((1, 1)), 12.0,
                         np.zeros((0, 3)), [], np.zeros((1, 3)), [1.0])
    assert s.rep != (1, 1, 1)
    for t in range(5):
        x = rng.random(3)
        sr1, lr1 = s.phi(x, form="eikr")
        sr2, lr2 = s.phi(x, form="sk")
        assert math.isclose(sr1, sr2, rel_tol=1e-13)
        assert abs(lr1 - lr2) < 1e-9 * (abs(lr1) + 1.0)
    assert s.nk_sk < s.wt.size  # only vectors divisible by rep survive


# ------------------------------------------------ insertion stream ----
def test_philox_known_answers(orc):
    """Random123 known-answer vectors for Philox4x32-10."""
    assert orc.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == \
        [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                             [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_insertion_uniforms(orc):
    u = np.array([orc.insertion_uniforms(42, g) for g in range(4000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    np.testing.assert_allclose(u.mean(axis=0), 0.5, atol=0.015)
    np.testing.assert_allclose(u.var(axis=0), 1 / 12, atol=0.005)
    assert not np.array_equal(orc.insertion_uniforms(42, 0), orc.insertion_uniforms(43, 0))


def test_arvo_rotation(orc, pm):
    """src/molecule.jl:233-244 and test/molecule.jl:226-288: orthogonal,
    det +1, uniform orientation."""
    rng = np.random.Generator(np.random.PCG64(3))
    z = []
    for _ in range(6000):
        u = rng.random(3)
        R = orc.arvo_rotation(u)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert math.isclose(np.linalg.det(R), 1.0, rel_tol=1e-13)
        np.testing.assert_allclose(R, pm.arvo_rotation_matrix(*u), atol=1e-15)
        z.append(R @ np.array([0.0, 0.0, 1.0]))
    z = np.array(z)
    np.testing.assert_allclose(z.mean(axis=0), 0.0, atol=0.03)
    np.testing.assert_allclose(z.T @ z / len(z), np.eye(3) / 3, atol=0.02)


# ----------------------------------------------- Widom + energy grid ----
def test_widom_ideal_gas_probe(orc, pm):
    """A probe with eps = 0 and no charge has E = 0 everywhere: every weight
    is 1 and the per-block K_H is 1/(R T) (ideal gas)."""
    x, ff = _xe_sbmof1(pm)
    ff0 = pm.LJForceField("zero", {**ff.pure_sigma}, {k: 0.0 for k in ff.pure_epsilon}, 12.5)
    s = _sys(orc, pm, x, ff0, ["Xe"], np.zeros((1, 3)))
    v, e = s.widom_energies(1, 0, 200)
    # overlaps still give Inf (the reference checks r^2 < 0.1 for any pair)
    fin = np.isfinite(v)
    assert np.all(v[fin] == 0.0) and np.all(e == 0.0)
    sums = s.block_sums(v, e, 298.0)
    assert sums[0] == fin.sum()
    r = pm.block_statistics(np.tile([[200.0, 0.0, 0.0]], (5, 1)), 200, 298.0, 1000.0)
    assert math.isclose(r["henry coefficient [mol/(m³-bar)]"], 1.0 / (pm.UNIV_GAS_CONST * 298.0),
                        rel_tol=1e-14)
    assert r["err henry coefficient [mmol/(g-bar)]"] == 0.0


def test_widom_single_pose_closed_form(pm):
    """All insertions with one energy E: K_H = e^{-E/T}/(R T), <U> = E."""
    E, T, n = -1234.5, 298.0, 1000
    w = math.exp(-E / T)
    sums = np.tile([[n * w, n * w * E, 0.0]], (5, 1))
    r = pm.block_statistics(sums, n, T, 800.0)
    assert math.isclose(r["henry coefficient [mol/(m³-bar)]"], w / (pm.UNIV_GAS_CONST * T), rel_tol=1e-13)
    assert math.isclose(r["⟨U⟩ (K)"], E, rel_tol=1e-13)
    assert math.isclose(r["Qst (kJ/mol)"], -E * pm.K_TO_KJ_PER_MOL + T * pm.K_TO_KJ_PER_MOL, rel_tol=1e-13)
    assert len(r) == 14  # plus "xtal" and "elapsed time (min)" in henry_coefficient


def test_widom_block_sums_semantics(orc):
    """src/henry.jl:236-243: Inf energies add 0 to S and are left out of W."""
    v = np.array([-100.0, np.inf, 50.0])
    e = np.array([-10.0, -5.0, np.inf])
    out = np.zeros(3)
    from oracle.oracle import lib, _d
    lib().orc_widom_block_sums(3, _d(v), _d(e), 300.0, _d(out))
    w0 = math.exp(110.0 / 300.0)
    assert math.isclose(out[0], w0, rel_tol=1e-15)
    assert math.isclose(out[1], -100.0 * w0, rel_tol=1e-15)
    assert math.isclose(out[2], -10.0 * w0, rel_tol=1e-15)


def test_widom_energies_match_direct_calls(orc, pm):
    """Each Widom energy equals vdw_energy / _phi at the pose the stream gives."""
    case = pm.synthetic.co2_widom_case(n_per_block=10)
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    lj_off, q_off = pm.site_offsets(co2)
    s = _sys(orc, pm, x, ff, co2.atom_species, lj_off.T, q_off.T, co2.charges_q)
    v, e = s.widom_energies(5, 100, 4)
    for i in range(4):
        u = orc.insertion_uniforms(5, 100 + i)
        R = orc.arvo_rotation(u[3:])
        sites = u[:3] + (s.c_to_f @ (R @ lj_off)).T
        assert math.isclose(s.vdw(sites), v[i], rel_tol=1e-13) or (math.isinf(v[i]) and math.isinf(s.vdw(sites)))
        es = sum(qc * sum(s.phi(p)) for qc, p in zip(co2.charges_q, sites))
        assert math.isclose(es, e[i], rel_tol=1e-12, abs_tol=1e-9)
    # reference e^{ik.r} form gives the same energies
    v2, e2 = s.widom_energies(5, 100, 2, es_mode=1)
    np.testing.assert_allclose(e2, e[:2], rtol=1e-10, atol=1e-8)


def test_energy_grid_faces_periodic(orc, pm):
    """test/grid.jl:43-55: with 0 and 1 both on the grid, opposite faces of
    an LTA CH4 grid agree (atol 1e-7 after clipping huge values)."""
    x = pm.read_cif(golden("crystals", "LTA.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    s = _sys(orc, pm, x, ff, ["CH4"], np.zeros((1, 3)))
    n_pts = pm.required_n_pts(x.box, 5.0)
    g = s.energy_grid(n_pts).reshape(n_pts, order="F")
    g[g > 1000.0] = 0.0
    np.testing.assert_allclose(g[0], g[-1], atol=1e-7)
    np.testing.assert_allclose(g[:, 0], g[:, -1], atol=1e-7)
    np.testing.assert_allclose(g[:, :, 0], g[:, :, -1], atol=1e-7)
    # grid point (i, j, k) is the probe at (i/(n-1), j/(n-1), k/(n-1)) ./ rep
    i, j, k = 1, 2, 3
    xf = np.array([i / (n_pts[0] - 1), j / (n_pts[1] - 1), k / (n_pts[2] - 1)]) / np.array(s.rep)
    e = s.vdw(xf.reshape(1, 3))
    assert math.isclose(s.energy_grid(n_pts)[i + n_pts[0] * (j + n_pts[1] * k)], e, rel_tol=1e-15) \
        or (math.isinf(e))
