"""Pose forces (pm_pose_gradient) and the energy minimiser built on them.

* Analytic dE/d(site xyz) of 64 random CO2 poses in the CO2 workload's
  framework (LJ + Ewald) vs central finite differences (h = 1e-4 A) of the
  oracle's FP64 energies of the displaced sites, to 1e-4 of the site's
  gradient norm + 1 K/A; the mixed path's forces vs FP64 to the same bar.
* find_energy_minimum: the reference's documented minimum of Xe in SBMOF-1
  with UFF (docs/src/energy_min.md: grid search at 1 A, then local
  minimisation; -37.69376112588296 kJ/mol at xf = (0.0175, 0.9373, 0.0112)):
  the energy there to 1e-12, the minimiser's end point near it.
* The same forces against the oracle's dual-number derivatives
  (oracle/dual.c): FP64 to 1e-8, mixed to 1e-4 of |g_site| + 1 K/A.
* find_energy_minima: 256 CO2 start poses minimised together."""
import math
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def co2sys(pm, orc):
    case = pm.synthetic.co2_widom_case()
    x, ff = case.crystal, case.ljff
    co2 = case.probes["CO2"]
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision, verbose=False)
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                         [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table, ff.sigma2_table,
                         math.sqrt(ff.r2_cutoff), lj_off.T, [ff.index(s) for s in co2.atom_species], q_off.T,
                         co2.charges_q, ewald_precision=case.ewald_precision)
    # CO2's LJ spheres and charges coincide: 3 sites with both
    species = [ff.index(s) for s in co2.atom_species]
    charges = list(co2.charges_q)
    return case, xr, fw, ew, o, lj_off.T, species, charges


def _random_poses(xr, offsets, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    poses = []
    for _ in range(n):
        com = rng.random(3) @ xr.box.f_to_c.T
        from porousmaterials_jl_amd import arvo_rotation_matrix
        R = arvo_rotation_matrix(*rng.random(3))
        sites = com[None, :] + offsets @ R.T
        poses.append(sites @ xr.box.c_to_f.T)
    return np.array(poses)  # (n, 3, 3) replicated fractional


def _oracle_energy(o, xr, sites_cart, charges):
    xf = sites_cart @ xr.box.c_to_f.T
    e = o.vdw(xf)
    for s, q in enumerate(charges):
        sr, lr = o.phi(xf[s])
        e += q * (sr + lr)
    return e


def test_pose_forces_match_oracle_finite_differences(pm, co2sys):
    say("case: pm_pose_gradient, 64 CO2 poses, FP64 and mixed vs central FD of oracle energies")
    case, xr, fw, ew, o, off, species, charges = co2sys
    poses = _random_poses(xr, off, 64, 99)
    e64, g64 = pm.device.pose_gradients(fw, case.ljff, ew, species, charges, poses, "fp64")
    e32, g32 = pm.device.pose_gradients(fw, case.ljff, ew, species, charges, poses, "mixed")
    h = 1e-4
    worst64 = worst32 = 0.0
    checked = straddled = 0
    for p in range(poses.shape[0]):
        if not np.isfinite(e64[p].sum()):
            continue
        cart = poses[p] @ xr.box.f_to_c.T
        want = _oracle_energy(o, xr, cart, charges)
        assert abs(e64[p].sum() - want) <= 1e-9 * (abs(want) + 1e3)
        def fd(s, d, step):
            cp, cm = cart.copy(), cart.copy()
            cp[s, d] += step
            cm[s, d] -= step
            return (_oracle_energy(o, xr, cp, charges) - _oracle_energy(o, xr, cm, charges)) / (2 * step)

        for s in range(3):
            tol = np.linalg.norm(g64[p, s]) + 1.0
            for d in range(3):
                a, b = fd(s, d, h), fd(s, d, h / 2)
                if abs(a - b) > 1e-5 * tol:
                    # the step straddles a cutoff sphere: the truncated LJ / erfc
                    # energy jumps there, so the difference quotient is no derivative
                    straddled += 1
                    continue
                worst64 = max(worst64, abs(g64[p, s, d] - b) / tol)
                checked += 1
            worst32 = max(worst32, np.max(np.abs(g32[p, s] - g64[p, s])) / tol)
    say(f"   {checked} gradient components checked ({straddled} FD steps straddled a cutoff): FP64 vs FD max "
        f"{worst64:.2e}, mixed vs FP64 max {worst32:.2e} (of |g_site| + 1 K/A)")
    assert checked >= 400 and straddled <= 0.05 * (checked + straddled)
    assert worst64 <= 1e-4
    assert worst32 <= 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pose_forces_match_dual_oracle(pm, co2sys, precision):
    """dE/d(site xyz) of 256 CO2 poses vs the oracle's dual-number forces
    (oracle/dual.c: the same energy with dual site coordinates, exact
    derivatives): FP64 to 1e-8 and mixed to 1e-4 of |g_site| + 1 K/A, the
    energies to 1e-10 / 1e-5 of |E| + 1e3 K."""
    say(f"case: pm_pose_gradient, 256 CO2 poses vs the dual-number oracle ({precision})")
    case, xr, fw, ew, o, off, species, charges = co2sys
    poses = _random_poses(xr, off, 256, 123)
    e, g = pm.device.pose_gradients(fw, case.ljff, ew, species, charges, poses, precision)
    ed, gd = o.pose_grad(species, charges, poses)
    et = e.sum(axis=1)
    assert np.array_equal(np.isinf(et), np.isinf(ed))
    fin = np.isfinite(ed)
    assert fin.sum() >= 32
    tol_e, tol_g = (1e-10, 1e-8) if precision == "fp64" else (1e-5, 1e-4)
    de = np.abs(et[fin] - ed[fin]) / (np.abs(ed[fin]) + 1e3)
    scale = np.linalg.norm(gd[fin], axis=2, keepdims=True) + 1.0
    dg = np.abs(g[fin] - gd[fin]) / scale
    say(f"   {fin.sum()} finite poses: energies {de.max():.2e} of |E| + 1e3 K, forces {dg.max():.2e} of "
        f"|g_site| + 1 K/A")
    assert de.max() <= tol_e and dg.max() <= tol_g


def test_find_energy_minimum_reference_doc_example(pm):
    """docs/src/energy_min.md: Xe in SBMOF-1, UFF (r_c 14 A): grid search at
    1 A resolution (13 x 7 x 24 points), then the local minimum.

    The LJ potential is truncated, not shifted, so the energy jumps where an
    atom crosses a cutoff sphere: the printed minimum lies on the low side of
    such a step, 0.016 A from the stationary point of the smooth part where
    a gradient method stops.  Pinned: the device energy at the printed
    minimum (-37.69376112588296 kJ/mol, to 1e-12), and the minimiser's end
    point (L-BFGS, then the shell search and Nelder-Mead polish) within
    0.05 A of it (or of its mirror image) at an energy within 1e-6 of it."""
    say("case: find_energy_minimum_gridsearch + find_energy_minimum, Xe / SBMOF-1 / UFF")
    x = pm.read_cif(golden("crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(golden("forcefields", "UFF.csv"))
    xe = pm.Molecule.from_dir(golden("molecules", "Xe"))
    ref_xf = np.array([0.01749943805846959, 0.9372916114895011, 0.011192272400742498])
    ref_e = -37.69376112588296
    rep = np.array(pm.replication_factors(x.box, ff))
    xr = pm.replicate(x, tuple(rep))
    fw = pm.device.framework(xr, ff)
    e_ref, _ = pm.device.pose_gradients(fw, ff, None, [ff.index("Xe")], [0.0], (ref_xf / rep).reshape(1, 1, 3),
                                        "fp64")
    assert abs(e_ref.sum() * 8.314 / 1000.0 - ref_e) <= 1e-12 * abs(ref_e)
    m0, e0 = pm.find_energy_minimum_gridsearch(x, xe, ff, resolution=1.0)
    assert pm.required_n_pts(x.box, 1.0) == (13, 7, 24)
    m1, e1 = pm.find_energy_minimum(x, m0, ff)
    # the grid search ends on SBMOF-1's inversion centre (a saddle, zero
    # gradient), from which the minimum and its mirror image -xf (the same
    # energy) are equally near; rounding decides which side a descent takes
    dist = np.inf
    for target in (ref_xf, -ref_xf):
        d = np.asarray(m1.com) - target
        d -= np.round(d)
        dist = min(dist, float(np.linalg.norm(x.box.f_to_c @ d)))
    say(f"   grid minimum {e0:.6f} kJ/mol at {m0.com}; minimised {e1:.12f} kJ/mol at {m1.com}, "
        f"{dist:.4f} A from the reference's minimum ({ref_e}) or its mirror image")
    assert e1 <= e0 + 1e-12 * abs(e0)
    assert dist <= 0.05
    # the reference's Nelder-Mead ends on the low side of a cutoff step (an
    # atom 14 A away entering the cutoff sphere lowers E by ~1e-3 kJ/mol), a
    # few thousandths of an A from the stationary point of the smooth part;
    # the shell search and Nelder-Mead polish after the L-BFGS reach it
    assert abs(e1 - ref_e) <= 1e-6 * abs(ref_e), (e1, ref_e)


def test_find_energy_minima_batched_co2(pm, co2sys):
    say("case: find_energy_minima, 256 CO2 start poses (LJ + Ewald forces), FP64")
    case, xr, fw, ew, o, off, species, charges = co2sys
    rng = np.random.Generator(np.random.PCG64(5))
    starts = rng.random((256, 3))
    co2 = case.probes["CO2"]
    xf, e, info = pm.find_energy_minima(case.crystal, co2, case.ljff, starts, max_iter=300)
    # start energies of the same poses (unit-cell fractional COM / rep)
    rep = np.array(pm.replication_factors(case.crystal.box, case.ljff))
    sites0 = (starts / rep) @ xr.box.f_to_c.T
    poses0 = (sites0[:, None, :] + off[None, :, :]) @ xr.box.c_to_f.T
    e0, _ = pm.device.pose_gradients(fw, case.ljff, ew, species, charges, poses0, "fp64")
    e0 = e0.sum(axis=1) * 8.314 / 1000.0
    fin = np.isfinite(e0)
    conv = info["converged"]
    say(f"   {info['iterations']} iterations, {info['device_calls']} device calls; {conv.sum()}/256 converged "
        f"to |g| < 1e-4 K/A; mean E {np.mean(e0[fin]):.1f} -> {np.mean(e[fin]):.1f} kJ/mol")
    assert np.all(e[fin] <= e0[fin] + 1e-9)
    assert conv.sum() >= 0.9 * 256
    assert np.all((xf >= 0.0) & (xf < 1.0))
