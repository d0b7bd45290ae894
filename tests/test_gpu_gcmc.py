"""Grand-canonical Monte Carlo chains on the device (pm_gcmc_*) against the
oracle's chain (oracle/gcmc.c) and the reference's known answers.

Parity: device chains and oracle chains draw the same random stream, so they
make the same proposals; with FP64 energies on both sides (different
summation order) they take the same decisions.  After a run the counts,
proposal/acceptance counters, block statistics, step size and density
counts must be equal, the energies equal to 1e-9 (relative) and the
coordinates to 1e-9.  Cases: Xe in SBMOF-1 (one site), CO2 in a synthetic
framework (three sites, rotations), Kr/Xe in SBMOF-1 (identity changes) with
a density grid, and a chain started from a given configuration.
"""
import math
import sys
import time

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_chains(ogc, pm, x, ff, templates, temps, fugs, n_burn, n_sample, seed, capacity, density=None,
                   rep=None):
    tmpl = []
    for m in templates:
        off = (m.atoms_x - m.com[:, None]).T
        tmpl.append((off, [ff.index(s) for s in m.atom_species], pm.needs_rotations(m)))
    out = []
    for c, (T, f) in enumerate(zip(temps, fugs)):
        out.append(ogc.OracleChain((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                                   [ff.index(s) for s in x.species], ff.epsilon_table, ff.sigma2_table,
                                   math.sqrt(ff.r2_cutoff), tmpl, T, f, n_burn, n_sample, seed, chain=c,
                                   capacity=capacity, density=density, rep=rep))
    return out


def _compare(dev, ochains, n_species, label):
    for c, oc in enumerate(ochains):
        d = dev.state(c)
        o = oc.state
        assert d["error"] == o.error == 0, label
        assert d["cycles_done"] == o.outer and d["markov_time"] == o.t, (label, c)
        assert list(d["n"][:n_species]) == list(o.n[:n_species]), (label, c, d["n"], list(o.n))
        assert list(d["n_proposed"]) == list(o.n_prop), (label, c)
        assert list(d["n_accepted"]) == list(o.n_acc), (label, c, list(d["n_accepted"]), list(o.n_acc))
        assert d["block"] == o.block and d["num_snapshots"] == o.num_snapshots
        assert list(d["block_samples"]) == list(o.blk_samples)
        assert np.array_equal(d["block_n"], np.array([list(r) for r in o.blk_n]))
        assert np.array_equal(d["block_n2"], np.array([list(r) for r in o.blk_n2]))
        assert d["delta"] == o.delta
        for a, b in ((d["energy_gh"], o.e_gh), (d["energy_gg"], o.e_gg)):
            assert abs(a - b) <= 1e-9 * (abs(b) + 1.0), (label, c, a, b)
        for key, ok in (("block_gh", o.blk_gh), ("block_gg", o.blk_gg), ("block_un", o.blk_un)):
            want = np.array(list(ok))
            assert np.allclose(d[key], want, rtol=1e-9, atol=1e-6), (label, key)
        for s in range(n_species):
            ds, dc = dev.molecules(c, s)
            os_, oc_ = oc.molecules(s)
            assert ds.shape == os_.shape
            assert np.max(np.abs(ds - os_), initial=0.0) <= 1e-9 and np.max(np.abs(dc - oc_), initial=0.0) <= 1e-9


def test_gcmc_xe_sbmof1_matches_oracle(pm, orc):
    from oracle import gcmc as ogc
    say("case: GCMC Xe/SBMOF-1 Dreiding, 3 chains (20, 200, 2000 Pa), 300 burn + 100 sample cycles vs oracle")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    rep = pm.replication_factors(x.box, ff)
    fw = pm.device.framework(pm.replicate(x, rep), ff)
    fugs = [[20.0], [200.0], [2000.0]]
    dev = pm.device.GcmcChains(fw, ff, [xe], [298.0] * 3, fugs, 300, 100, capacity=256, seed=17)
    t = time.time()
    dev.run(170)
    dev.run(230)
    dt = time.time() - t
    och = _oracle_chains(ogc, pm, x, ff, [xe], [298.0] * 3, fugs, 300, 100, 17, 256)
    for oc in och:
        oc.run(400)
    moves = sum(dev.state(c)["markov_time"] for c in range(3))
    say(f"   {moves} moves over 3 chains in {dt:.2f} s; N = {[dev.state(c)['n'][0] for c in range(3)]}")
    _compare(dev, och, 1, "Xe")
    for c in range(3):
        gh, gg = dev.total_energy(c)
        st = dev.state(c)
        assert abs(gh - st["energy_gh"]) <= 1e-6 * (abs(gh) + 1) and abs(gg - st["energy_gg"]) <= 1e-6 * (abs(gg) + 1)


def test_gcmc_co2_rotations_match_oracle(pm, orc):
    from oracle import gcmc as ogc
    say("case: GCMC CO2 (3 sites) in the 424-atom synthetic framework, 2 chains, 100 burn + 50 sample cycles")
    case = pm.synthetic.lj_grid_case()
    x, ff = case.crystal, case.ljff
    co2 = pm.synthetic.co2_molecule()
    rep = pm.replication_factors(x.box, ff)
    fw = pm.device.framework(pm.replicate(x, rep), ff)
    fugs = [[2e6], [2e7]]
    dev = pm.device.GcmcChains(fw, ff, [co2], [260.0, 300.0], fugs, 100, 50, capacity=512, seed=3)
    dev.run(150)
    och = _oracle_chains(ogc, pm, x, ff, [co2], [260.0, 300.0], fugs, 100, 50, 3, 512)
    for oc in och:
        oc.run(150)
    st = [dev.state(c) for c in range(2)]
    say(f"   N = {[s['n'][0] for s in st]}, rotations accepted {[int(s['n_accepted'][3]) for s in st]}")
    assert all(s["n_proposed"][3] > 0 for s in st)
    _compare(dev, och, 1, "CO2")


def test_gcmc_mixture_identity_density_match_oracle(pm, orc):
    from oracle import gcmc as ogc
    say("case: GCMC Kr/Xe (0.9/0.1 of 0.5 bar) in SBMOF-1, UFF, identity changes and a Xe density grid")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    kr = pm.Molecule.from_dir(golden("molecules", "Kr"))
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    fw = pm.device.framework(xr, ff)
    fugs = [[0.45e5, 0.05e5], [0.09e5, 0.01e5]]
    dens = {"species": 1, "site_mask": 1, "n_pts": (9, 11, 13), "sim_box": True}
    dev = pm.device.GcmcChains(fw, ff, [kr, xe], [298.0] * 2, fugs, 200, 50, capacity=256, seed=11, density=dens)
    dev.run(60)
    dev.run(190)
    och = _oracle_chains(ogc, pm, x, ff, [kr, xe], [298.0] * 2, fugs, 200, 50, 11, 256, density=dens)
    for oc in och:
        oc.run(250)
    st = [dev.state(c) for c in range(2)]
    say(f"   N(Kr, Xe) = {[list(s['n'][:2]) for s in st]}, identity changes accepted "
        f"{[int(s['n_accepted'][5]) for s in st]}")
    assert all(s["n_proposed"][5] > 0 for s in st)
    _compare(dev, och, 2, "Kr/Xe")
    for c, oc in enumerate(och):
        assert np.array_equal(dev.density(c), oc.density())
    assert dev.density(0).sum() > 0


def test_gcmc_start_configuration_matches_oracle(pm, orc):
    from oracle import gcmc as ogc
    say("case: GCMC Xe/SBMOF-1 started from a given configuration (energy recomputed), 80 cycles")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    rep = pm.replication_factors(x.box, ff)
    fw = pm.device.framework(pm.replicate(x, rep), ff)
    dev0 = pm.device.GcmcChains(fw, ff, [xe], [298.0], [[2000.0]], 400, 5, capacity=256, seed=17)
    dev0.run(400)
    sites, com = dev0.molecules(0, 0)
    assert sites.shape[0] > 5
    dev = pm.device.GcmcChains(fw, ff, [xe], [298.0], [[500.0]], 40, 40, capacity=256, seed=2)
    dev.set_molecules(0, 0, sites, com)
    dev.run(80)
    oc = _oracle_chains(ogc, pm, x, ff, [xe], [298.0], [[500.0]], 40, 40, 2, 256)[0]
    oc.set_molecules(0, sites, com)
    oc.run(80)
    _compare(dev, [oc], 1, "restart")


def _states(dev, n_chains, n_species):
    out = []
    for c in range(n_chains):
        st = dev.state(c)
        mols = [dev.molecules(c, s) for s in range(n_species)]
        out.append((st, mols))
    return out


def test_gcmc_molecule_store_placement_is_invisible(pm, orc):
    """The chains keep their molecules in LDS when they fit and move them to
    device memory when a chain could outgrow the LDS store; a 4-molecule
    store makes every chain overflow in mid-call.  Every placement gives the
    same chains bit for bit, and the oracle's."""
    from oracle import gcmc as ogc
    say("case: GCMC Xe/SBMOF-1 Dreiding 3 chains and Kr/Xe UFF 2 chains with the molecule store in LDS "
        "(default, 4 per species) and in device memory")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    cases = []
    ff = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    cases.append(("Xe", ff, [xe], [298.0] * 3, [[20.0], [200.0], [2000.0]], 300, 100, 17, None))
    uff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    kr = pm.Molecule.from_dir(golden("molecules", "Kr"))
    xe_u = pm.Molecule.from_dir(golden("molecules", "Xe"))
    dens = {"species": 1, "site_mask": 1, "n_pts": (9, 11, 13), "sim_box": True}
    cases.append(("Kr/Xe", uff, [kr, xe_u], [298.0] * 2, [[0.45e5, 0.05e5], [0.09e5, 0.01e5]], 200, 50, 11, dens))
    for label, f, tmpl, temps, fugs, nb, ns, seed, density in cases:
        rep = pm.replication_factors(x.box, f)
        fw = pm.device.framework(pm.replicate(x, rep), f)
        runs = {}
        for limit in (-1, 4, 0):
            dev = pm.device.GcmcChains(fw, f, tmpl, temps, fugs, nb, ns, capacity=256, seed=seed, density=density)
            dev.set_lds_molecules(limit)
            dev.run(150)
            dev.run(nb + ns - 150)
            runs[limit] = (dev, _states(dev, len(temps), len(tmpl)))
        base = runs[-1][1]
        for limit in (4, 0):
            other = runs[limit][1]
            for (sa, ma), (sb, mb) in zip(base, other):
                for key in sa:
                    assert np.array_equal(sa[key], sb[key]), (label, limit, key)
                for (a_s, a_c), (b_s, b_c) in zip(ma, mb):
                    assert np.array_equal(a_s, b_s) and np.array_equal(a_c, b_c), (label, limit)
            if density is not None:
                for c in range(len(temps)):
                    assert np.array_equal(runs[-1][0].density(c), runs[limit][0].density(c))
        # the 4-molecule store overflowed: some chain ended with n_s + N + 1 > 4
        assert max(int(st["n"].max()) + int(st["n"].sum()) + 1 for st, _ in base) > 4
        och = _oracle_chains(ogc, pm, x, f, tmpl, temps, fugs, nb, ns, seed, 256, density=density)
        for oc in och:
            oc.run(nb + ns)
        _compare(runs[4][0], och, len(tmpl), label + " (4-molecule LDS store)")
        say(f"   {label}: N = {[list(s['n'][:len(tmpl)]) for s, _ in base]}, placements agree bit for bit")


def _gg_box(pm, length, ff, templates):
    """An empty cubic box and a GCMC handle whose from-scratch energy
    (pm_gcmc_total_energy) is the guest-guest energy of the molecules set."""
    box = pm.Box(length, length, length, math.pi / 2, math.pi / 2, math.pi / 2)
    x = pm.Crystal("empty", box, [], np.zeros((3, 0)))
    dev = pm.device.GcmcChains(pm.device.framework(x, ff), ff, templates, [298.0], [[1.0] * len(templates)], 0, 5,
                               capacity=8)
    return box, dev


def _placed(box, tmpl, cart_coms):
    """translate_to! of unrotated templates: sites (n, n_sites, 3) and centres
    (n, 3), fractional."""
    off = tmpl.atoms_x - tmpl.com[:, None]
    sites = np.array([(box.c_to_f @ (np.asarray(c, dtype=float)[:, None] + off)).T for c in cart_coms])
    coms = np.array([box.c_to_f @ np.asarray(c, dtype=float) for c in cart_coms])
    return sites, coms


def test_gcmc_guest_guest_energy_known_answers(pm):
    """The device's guest-guest energy against test/vdw_energetics.jl:97-227:
    Xe-He at 6 A directly and through the periodic boundary, a molecule on top
    of another (+Inf), a three-axis periodic pair, the cutoff, two CO2 6 A
    apart (directly and through the boundary), CO2 pairs with only some sites
    inside the cutoff, and a Xe/Xe/Kr mixture.  The total counts every guest
    pair once."""
    say("case: guest-guest energies of Xe/He, CO2/CO2 and Xe/Xe/Kr in empty boxes (the reference's closed forms)")
    dre = pm.LJForceField.from_csv(golden("forcefields", "Dreiding.csv"), r_cutoff=12.5)
    xe, he = pm.lj_sphere("Xe"), pm.lj_sphere("He")
    co2 = pm.Molecule.from_dir(golden("molecules", "CO2"))

    def lj(ff, r2, a, b):
        i, j = ff.index(a), ff.index(b)
        return pm.lennard_jones(r2, ff.sigma2_table[i, j], ff.epsilon_table[i, j])

    def gg(dev):
        gh, g = dev.total_energy(0)
        assert gh == 0.0
        return g

    box, dev = _gg_box(pm, 25.0, dre, [xe, he])
    for a, b, r2 in (([5.0, 12.0, 12.0], [11.0, 12.0, 12.0], 36.0), ([5.0, 12.0, 12.0], [24.0, 12.0, 12.0], 36.0),
                     ([24.0, 23.0, 11.0], [22.0, 2.0, 12.0], 21.0), ([11.0, 1.0, 12.0], [11.0, 4.0, 12.0], 9.0),
                     ([11.0, 1.0, 12.0], [11.0, 23.0, 12.0], 9.0)):
        dev.set_molecules(0, 0, *_placed(box, xe, [a]))
        dev.set_molecules(0, 1, *_placed(box, he, [b]))
        assert gg(dev) == pytest.approx(lj(dre, r2, "Xe", "He"), rel=1e-12)
    dev.set_molecules(0, 0, *_placed(box, xe, [[0.0, 0.0, 0.0]]))
    dev.set_molecules(0, 1, *_placed(box, he, [[12.0, 12.0, 12.0]]))
    assert gg(dev) == 0.0  # beyond the cutoff
    dev.set_molecules(0, 0, *_placed(box, xe, [[5.0, 12.0, 12.0], [5.0, 12.0, 12.0]]))
    dev.set_molecules(0, 1, *_placed(box, he, [[24.0, 12.0, 12.0]]))
    assert gg(dev) == math.inf  # a molecule on top of another

    box, dev = _gg_box(pm, 25.0, dre, [co2])
    r2_co, r2_oo = 1.16 ** 2, (2.0 * 1.16) ** 2
    for a, b in (([12.0, 9.0, 12.0], [12.0, 15.0, 12.0]), ([12.0, 2.0, 12.0], [12.0, 21.0, 12.0])):
        dev.set_molecules(0, 0, *_placed(box, co2, [a, b]))
        want = (2.0 * lj(dre, 36.0, "O_CO2", "O_CO2") + 4.0 * lj(dre, 36.0 + r2_co, "O_CO2", "C_CO2")
                + 2.0 * lj(dre, 36.0 + r2_oo, "O_CO2", "O_CO2") + lj(dre, 36.0, "C_CO2", "C_CO2"))
        assert gg(dev) == pytest.approx(want, rel=1e-12)
    box, dev = _gg_box(pm, 50.0, dre, [co2])
    dev.set_molecules(0, 0, *_placed(box, co2, [[0.0, 0.0, 0.0], [13.0, 0.0, 0.0]]))
    want = lj(dre, (13.0 - 2.0 * 1.16) ** 2, "O_CO2", "O_CO2") + 2.0 * lj(dre, (13.0 - 1.16) ** 2, "O_CO2", "C_CO2")
    assert gg(dev) == pytest.approx(want, rel=1e-12)

    uff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"), r_cutoff=6.0)
    box, dev = _gg_box(pm, 10.0, uff, [xe, pm.lj_sphere("Kr")])
    dev.set_molecules(0, 0, *_placed(box, xe, [[1.0, 1.0, 1.0], [1.0, 5.0, 1.0]]))
    dev.set_molecules(0, 1, *_placed(box, pm.lj_sphere("Kr"), [[6.0, 1.0, 1.0]]))
    # Xe1-Xe2 at 4 A, Xe1-Kr at 5 A, Xe2-Kr at sqrt(41) A > r_c
    assert gg(dev) == pytest.approx(lj(uff, 16.0, "Xe", "Xe") + lj(uff, 25.0, "Xe", "Kr"), rel=1e-12)
    dev.set_molecules(0, 0, *_placed(box, xe, [[1.0, 1.0, 1.0]]))
    dev.set_molecules(0, 1, np.zeros((0, 1, 3)), np.zeros((0, 3)))
    assert gg(dev) == 0.0
