"""The oracle's dual-number derivatives (oracle/dual.c) against the oracle
itself: the dual values equal the plain oracle's energies, and the dual
derivatives equal central finite differences of the plain oracle (the
mixing rule applied to perturbed pure parameters; perturbed charges;
displaced sites) to 1e-6 of the difference quotient plus its rounding
noise (64 ulp of |E| over the step), wherever the quotient is a derivative
(steps that carry an atom across a cutoff sphere are detected by a
half-step repeat and skipped, as the truncated potential jumps there).  A
planted sign error in the dual's derivatives (negated here, on the CPU)
fails the comparison."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def co2sys(pm, orc):
    case = pm.synthetic.co2_widom_case()
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    return case, x, ff, co2


def _oracle(orc, pm, x, ff, co2, q=None):
    lj_off, q_off = pm.site_offsets(co2)
    return orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                            [ff.index(s) for s in x.species], x.atom_q if q is None else q, ff.epsilon_table,
                            ff.sigma2_table, math.sqrt(ff.r2_cutoff), lj_off.T,
                            [ff.index(s) for s in co2.atom_species], q_off.T, co2.charges_q, ewald_precision=1e-6)


def _perturbed(pm, ff, species, which, factor):
    eps, sig = dict(ff.pure_epsilon), dict(ff.pure_sigma)
    (eps if which == "epsilon" else sig)[species] *= factor
    return pm.LJForceField(ff.name, sig, eps, math.sqrt(ff.r2_cutoff), ff.mixing_rules)


def _pure(ff):
    return (np.array([ff.pure_epsilon[s] for s in ff.species]), np.array([ff.pure_sigma[s] for s in ff.species]))


def test_dual_values_equal_the_oracle(pm, orc, co2sys):
    case, x, ff, co2 = co2sys
    o = _oracle(orc, pm, x, ff, co2)
    pe, ps = _pure(ff)
    ft = sorted({ff.index(s) for s in x.species})
    v, e, d = o.widom_param_grad(2024, 0, 600, pe, ps, ft)
    v0, e0 = o.widom_energies(2024, 0, 600)
    assert np.array_equal(np.isinf(v), np.isinf(v0))
    fin = np.isfinite(v0)
    assert np.max(np.abs(v[fin] - v0[fin]) / (np.abs(v0[fin]) + 1.0)) < 1e-12
    assert np.max(np.abs(e[fin] - e0[fin]) / (np.abs(e0[fin]) + 1.0)) < 1e-10
    assert d.shape == (600, len(ft), 3) and np.all(d[~fin] == 0.0)


def _check_param_derivatives(pm, orc, co2sys, negate=False):
    case, x, ff, co2 = co2sys
    o = _oracle(orc, pm, x, ff, co2)
    pe, ps = _pure(ff)
    ft = sorted({ff.index(s) for s in x.species})
    n = 300
    v, e, d = o.widom_param_grad(77, 0, n, pe, ps, ft)
    if negate:
        d = -d
    fin = np.isfinite(v)
    E = v + e
    q0 = np.array(x.atom_q, dtype=float)
    checked = 0
    worst = 0.0
    for i, t in enumerate(ft):
        sp = ff.species[t]
        for k, which in enumerate(("epsilon", "sigma", "q")):
            def energies(h):
                if which == "q":
                    mask = np.array([s == sp for s in x.species])
                    oo = _oracle(orc, pm, x, ff, co2, q0 + h * mask)
                else:
                    oo = _oracle(orc, pm, x, _perturbed(pm, ff, sp, which, 1.0 + h), co2)
                a, b = oo.widom_energies(77, 0, n)
                return a + b
            h = 1e-5
            scale = 1.0 if which == "q" else (ff.pure_epsilon[sp] if which == "epsilon" else ff.pure_sigma[sp])
            ep, em = energies(h / 2), energies(-h / 2)
            fd1 = (energies(h) - energies(-h)) / (2 * h * scale)
            fd2 = (ep - em) / (h * scale)
            # rounding of the energies themselves: a few ulp of |E| over the step
            with np.errstate(invalid="ignore"):
                noise = 64 * 2.2e-16 * np.maximum(np.abs(ep), np.abs(em)) / (h * scale)
                ok = fin & np.isfinite(fd2) & (np.abs(fd1 - fd2) <= 1e-6 * np.abs(fd2) + 2 * noise + 1e-6)
            bound = 1e-6 * np.abs(fd2[ok]) + noise[ok] + 1e-6
            ratio = np.abs(d[ok, i, k] - fd2[ok]) / bound
            worst = max(worst, float(ratio.max()))
            checked += int(ok.sum())
    return worst, checked


def test_dual_parameter_derivatives_match_finite_differences(pm, orc, co2sys):
    worst, checked = _check_param_derivatives(pm, orc, co2sys)
    assert checked > 1000
    assert worst <= 1.0, worst


def test_planted_sign_error_is_caught(pm, orc, co2sys):
    worst, _ = _check_param_derivatives(pm, orc, co2sys, negate=True)
    assert worst > 1e3


def test_dual_pose_forces_match_finite_differences(pm, orc, co2sys):
    case, x, ff, co2 = co2sys
    o = _oracle(orc, pm, x, ff, co2)
    lj_off, _ = pm.site_offsets(co2)
    rng = np.random.Generator(np.random.PCG64(3))
    F, C = o.F, o.c_to_f
    poses = []
    for _ in range(64):
        com = rng.random(3) @ F.T
        R = pm.arvo_rotation_matrix(*rng.random(3))
        poses.append((com[None, :] + lj_off.T @ R.T) @ C.T)
    poses = np.array(poses)
    species = [ff.index(s) for s in co2.atom_species]
    e, g = o.pose_grad(species, co2.charges_q, poses)
    h = 1e-5
    worst, checked = 0.0, 0
    for p in range(len(poses)):
        if not np.isfinite(e[p]):
            continue
        cart = poses[p] @ F.T
        want = o.vdw(poses[p]) + sum(q * sum(o.phi(poses[p][s])) for s, q in enumerate(co2.charges_q))
        assert abs(e[p] - want) <= 1e-10 * (abs(want) + 1e3)
        for s in range(3):
            for dd in range(3):
                def en(step):
                    c2 = cart.copy()
                    c2[s, dd] += step
                    x2 = c2 @ C.T
                    return o.vdw(x2) + sum(q * sum(o.phi(x2[k])) for k, q in enumerate(co2.charges_q))
                a = (en(h) - en(-h)) / (2 * h)
                b = (en(h / 2) - en(-h / 2)) / h
                tol = np.linalg.norm(g[p, s]) + 1.0
                if abs(a - b) > 1e-6 * tol:
                    continue
                worst = max(worst, abs(g[p, s, dd] - b) / tol)
                checked += 1
    assert checked > 300
    assert worst < 1e-5, worst
