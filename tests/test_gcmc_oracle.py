"""The oracle's GCMC chain (oracle/gcmc.c) against closed forms of the
grand-canonical ensemble, on the CPU:

  * an ideal gas (no framework, epsilon 0) has Poisson-distributed N with
    <N> = var(N) = f V / (k_B T) (test/gcmc_long.jl:24-34 checks the mean);
  * a two-species ideal gas with identity changes has <N_i> = f_i V / (k_B T)
    for each species, which fixes the identity-change acceptance rule
    (src/gcmc.jl:574-627) as well as insertion and deletion;
  * rigid-body moves keep every molecule's internal geometry and keep the
    centres of mass inside the unit cube (test/mc_helpers.jl);
  * the running energies equal a from-scratch recomputation after a run
    (src/gcmc.jl:729-741).

The device chain reproduces the oracle's chain move for move
(tests/test_gpu_gcmc.py), so these closed forms pin both.
"""
import math

import numpy as np
import pytest

from oracle import gcmc as ogc

KB = 1.38064852e7


def _ideal_gas_chain(fugs, n_burn, n_sample, seed, L=30.0, capacity=256):
    S = len(fugs)
    eps = np.zeros((S, S))
    sig2 = np.full((S, S), 1e-12)
    tmpl = [(np.zeros((1, 3)), [s], False) for s in range(S)]
    return ogc.OracleChain((L, L, L, math.pi / 2, math.pi / 2, math.pi / 2), np.zeros((0, 3)), [], eps, sig2,
                           14.0, tmpl, 298.0, fugs, n_burn, n_sample, seed, capacity=capacity, rep=(1, 1, 1))


def _moments(st, s):
    samples = float(sum(st.blk_samples))
    mean = sum(st.blk_n[b][s] for b in range(5)) / samples
    var = sum(st.blk_n2[b][s] for b in range(5)) / samples - mean ** 2
    return mean, var


def test_oracle_ideal_gas_poisson():
    n_ig = 5.0
    f = n_ig * KB * 298.0 / 27000.0
    ch = _ideal_gas_chain([f], 500, 20000, seed=1)
    ch.run(20500)
    assert ch.state.error == 0
    mean, var = _moments(ch.state, 0)
    assert mean == pytest.approx(n_ig, rel=0.03)
    assert var == pytest.approx(n_ig, rel=0.1)
    # no framework and epsilon 0: the energy stays exactly zero
    assert ch.state.e_gh == 0.0 and ch.state.e_gg == 0.0


def test_oracle_ideal_gas_mixture_identity_changes():
    n1, n2 = 4.0, 1.5
    f = [n * KB * 298.0 / 27000.0 for n in (n1, n2)]
    ch = _ideal_gas_chain(f, 500, 20000, seed=2)
    ch.run(20500)
    st = ch.state
    assert st.n_prop[5] > 0 and st.n_acc[5] > 0  # identity changes proposed and accepted
    for s, n in enumerate((n1, n2)):
        mean, var = _moments(st, s)
        assert mean == pytest.approx(n, rel=0.04)
        assert var == pytest.approx(n, rel=0.12)


def test_oracle_rigid_moves_and_energy_bookkeeping(pm):
    case = pm.synthetic.lj_grid_case()
    x, ff = case.crystal, case.ljff
    co2 = pm.synthetic.co2_molecule()
    off = (co2.atoms_x - co2.com[:, None]).T
    tmpl = [(off, [ff.index(s) for s in co2.atom_species], True)]
    rep = pm.replication_factors(x.box, ff)
    ch = ogc.OracleChain((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), tmpl, 300.0, [2e7], 50, 50, 9, capacity=512)
    ch.run(100)
    st = ch.state
    assert st.n[0] > 0 and st.n_acc[2] > 0 and st.n_acc[3] > 0
    sites, com = ch.molecules(0)
    F = ch.F
    for i in range(sites.shape[0]):
        assert np.all(com[i] >= 0.0) and np.all(com[i] < 1.0)
        d = (F @ (sites[i] - com[i]).T).T  # Cartesian offsets from the centre
        assert np.allclose(np.linalg.norm(d, axis=1), np.linalg.norm(off, axis=1), atol=1e-9)
        assert np.allclose(np.linalg.norm(d[1] - d[0]), np.linalg.norm(off[1] - off[0]), atol=1e-9)
    gh, gg = ch.total_energy()
    assert gh == pytest.approx(st.e_gh, rel=1e-9, abs=1e-6)
    assert gg == pytest.approx(st.e_gg, rel=1e-9, abs=1e-6)
    assert tuple(ch.rep) == tuple(rep)


def _brute_energies(F, fw_xf, fw_type, eps, sig2, r2_cut, mols):
    """Guest-host and guest-guest LJ energies by brute force: the nearest of
    the 27 periodic images in Cartesian space, each guest pair once
    (src/vdw_energetics.jl:25-47,99-114 with R2_OVERLAP = 0.1)."""
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=float)
    cart_shifts = shifts @ F.T

    def pair(xa, ta, xb, tb):
        d = (F @ (xb - xa)[:, None]).ravel()
        r2 = np.min(np.sum((d[None, :] + cart_shifts) ** 2, axis=1))
        if r2 >= r2_cut:
            return 0.0
        if r2 < 0.1:
            return math.inf
        x = (sig2[ta, tb] / r2) ** 3
        return 4.0 * eps[ta, tb] * x * (x - 1.0)

    gh = 0.0
    for sites, types in mols:
        for s, t in zip(sites, types):
            for a, ta in zip(fw_xf, fw_type):
                gh += pair(s, t, a, ta)
    gg = 0.0
    for i in range(len(mols)):
        for j in range(i + 1, len(mols)):
            for s, t in zip(*mols[i]):
                for s2, t2 in zip(*mols[j]):
                    gg += pair(s, t, s2, t2)
    return gh, gg


def test_oracle_energies_match_brute_force(pm):
    """The chain's running energies and its from-scratch total against a
    brute-force sum (nearest of 27 images, pairs once): a double-counted
    guest pair, a dropped image or a missed cutoff fails here."""
    case = pm.synthetic.lj_grid_case()
    x, ff = case.crystal, case.ljff
    co2 = pm.synthetic.co2_molecule()
    off = (co2.atoms_x - co2.com[:, None]).T
    types = [ff.index(s) for s in co2.atom_species]
    tmpl = [(off, types, True), (np.zeros((1, 3)), [ff.index(co2.atom_species[0])], False)]
    ch = ogc.OracleChain((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), tmpl, 300.0, [2e7, 5e6], 300, 10, 13, capacity=512)
    ch.run(310)
    st = ch.state
    assert st.n[0] >= 3 and st.n[1] >= 3
    mols = []
    for s, ty in ((0, types), (1, tmpl[1][1])):
        sites, _ = ch.molecules(s)
        mols += [(sites[i], ty) for i in range(sites.shape[0])]
    gh, gg = _brute_energies(ch.F, ch.xf, ch.types, ch.eps, ch.sig2, ch.cfg.r2_cut, mols)
    assert gg != 0.0
    assert st.e_gh == pytest.approx(gh, rel=1e-9, abs=1e-6)
    assert st.e_gg == pytest.approx(gg, rel=1e-9, abs=1e-6)
    tgh, tgg = ch.total_energy()
    assert tgh == pytest.approx(gh, rel=1e-9, abs=1e-6) and tgg == pytest.approx(gg, rel=1e-9, abs=1e-6)
