"""Minimum-energy position of a rigid guest, on device forces.

Mirrors ``find_energy_minimum`` (``src/energy_min.jl:19-58``) and
``find_energy_minimum_gridsearch`` (``:75-85``): the guest-host energy of a
rigid molecule is minimised over its centre of mass (the orientation stays
as given, as in the reference), energies are returned in kJ/mol
(x 8.314/1000, ``:48``), the molecule comes back in unit-cell fractional
coordinates with its centre of mass wrapped into [0, 1) (``:51-54``).

Differences, by design:

* the reference runs Nelder-Mead (Optim) on ``vdw_energy``; here an L-BFGS
  iteration uses the analytic forces of ``pm_pose_gradient``, and many start
  poses advance together (``find_energy_minima``: one device call per
  iteration for all of them), followed by a batched Nelder-Mead polish of
  every end point, which finds the lower energies that lie across the
  cutoff's energy steps;
* when both the crystal and the molecule carry charges the Ewald energy and
  its forces are included (the reference raises an error there).
"""
from __future__ import annotations

import math

import numpy as np

from . import device
from .crystal import has_charges as xtal_has_charges, replicate
from .ewald import setup_Ewald_sum
from .forcefield import replication_factors
from .molecule import Frac, has_charges as mol_has_charges, needs_rotations, translate_to

K_TO_KJ = 8.314 / 1000.0  # src/energy_min.jl:48


class _Objective:
    """Total energy (K) and COM gradient (K/A) of P rigid poses of one
    molecule in a replicated crystal."""

    def __init__(self, xtal_rep, molecule_cart, ljff, ewald_precision, precision):
        self.box = xtal_rep.box
        charged = xtal_has_charges(xtal_rep) and mol_has_charges(molecule_cart)
        self.fw = device.framework(xtal_rep, ljff)
        self.ff = device.forcefield(ljff)
        self.ew = None
        if charged:
            ep = setup_Ewald_sum(xtal_rep.box, math.sqrt(ljff.r2_cutoff), eps=ewald_precision, verbose=False)
            self.ew = device.ewald(xtal_rep, self.fw, ep)
        m = molecule_cart
        lj_off = m.atoms_x - m.com[:, None]
        q_off = m.charges_x - m.com[:, None] if charged else np.zeros((3, 0))
        self.offsets = np.concatenate([lj_off, q_off], axis=1).T  # (n_sites, 3) Cartesian
        self.species = [ljff.index(s) for s in m.atom_species] + [-1] * q_off.shape[1]
        self.charges = [0.0] * m.n_atoms + (list(m.charges_q) if charged else [])
        self.precision = precision
        self.n_evals = 0

    def __call__(self, com_cart, idx=None):
        """com_cart: (P, 3) Cartesian centres of mass (``idx``, the poses'
        indices, does not matter: every pose has the same energy function)."""
        sites = com_cart[:, None, :] + self.offsets[None, :, :]
        xf = sites @ self.box.c_to_f.T
        e, g = device.pose_gradients(self.fw, self.ff, self.ew, self.species, self.charges, xf, self.precision)
        self.n_evals += 1
        return e.sum(axis=1), g.sum(axis=1)


def _lbfgs(obj, x0, m=10, max_iter=500, gtol=1e-4, max_step=0.5):
    """Batched L-BFGS with backtracking (Armijo) over P independent
    3-dimensional problems; ``obj(x, idx)`` evaluates the problems ``idx`` at
    ``x`` (rows).  Returns (x, f, g, iterations, converged)."""
    x = np.array(x0, dtype=np.float64)
    P = x.shape[0]
    f, g = obj(x, np.arange(P))
    S = np.zeros((m, P, 3))
    Y = np.zeros((m, P, 3))
    R = np.zeros((m, P))
    n_hist = np.zeros(P, dtype=int)
    head = 0
    gam = np.ones(P)  # initial inverse-Hessian scale s.y / y.y of each pose's latest pair
    done = ~np.isfinite(f) | (np.linalg.norm(g, axis=1) < gtol)
    it = 0
    for it in range(1, max_iter + 1):
        if done.all():
            break
        # two-loop recursion over the ring buffer (entries older than a pose's
        # history are masked out by R = 0)
        q = g.copy()
        alpha = np.zeros((m, P))
        order = [(head - 1 - i) % m for i in range(m)]
        for i in order:
            alpha[i] = R[i] * np.einsum("pk,pk->p", S[i], q)
            q -= alpha[i][:, None] * Y[i]
        r = np.where(n_hist > 0, gam, 1.0)[:, None] * q
        for i in reversed(order):
            b = R[i] * np.einsum("pk,pk->p", Y[i], r)
            r += S[i] * (alpha[i] - b)[:, None]
        d = -r
        gd = np.einsum("pk,pk->p", g, d)
        bad = ~(gd < 0)
        d[bad] = -g[bad]
        gd[bad] = -np.einsum("pk,pk->p", g[bad], g[bad])
        # first step of a fresh history: at most 0.1 A; never more than max_step
        dn = np.linalg.norm(d, axis=1)
        t = np.where(n_hist == 0, np.minimum(1.0, 0.1 / np.maximum(dn, 1e-300)), 1.0)
        t = np.minimum(t, max_step / np.maximum(dn, 1e-300))
        accepted = done.copy()
        x_new, f_new, g_new = x.copy(), f.copy(), g.copy()
        for _ in range(40):
            trial = ~accepted
            if not trial.any():
                break
            # only the poses still searching are evaluated (poses are independent)
            idx = np.nonzero(trial)[0]
            xt = x[idx] + t[idx, None] * d[idx]
            ft, gt = obj(xt, idx)
            # Armijo, with slack for the rounding of f near the minimum
            ok_s = np.isfinite(ft) & (ft <= f[idx] + 1e-4 * t[idx] * gd[idx] + 1e-14 * np.abs(f[idx]))
            ok = np.zeros(P, dtype=bool)
            ok[idx[ok_s]] = True
            x_new[idx[ok_s]], f_new[idx[ok_s]], g_new[idx[ok_s]] = xt[ok_s], ft[ok_s], gt[ok_s]
            accepted |= ok
            t = np.where(accepted, t, 0.5 * t)
        stuck = ~accepted
        moved = accepted & ~done
        s = x_new - x
        y = g_new - g
        sy = np.einsum("pk,pk->p", s, y)
        keep = moved & (sy > 1e-12 * np.linalg.norm(s, axis=1) * np.linalg.norm(y, axis=1))
        S[head][keep], Y[head][keep] = s[keep], y[keep]
        R[head] = np.where(keep, 1.0 / np.where(keep, sy, 1.0), 0.0)
        yy = np.einsum("pk,pk->p", y, y)
        gam = np.where(keep, sy / np.where(keep, yy, 1.0), gam)
        n_hist = np.where(keep, np.minimum(n_hist + 1, m), n_hist)
        head = (head + 1) % m
        x, f, g = x_new, f_new, g_new
        gn = np.linalg.norm(g, axis=1)
        # a pose whose line search failed restarts from steepest descent once;
        # failing again with a small gradient ends it
        restart = stuck & (n_hist > 0)
        n_hist[restart] = 0
        R[:, restart] = 0.0
        done |= (gn < gtol) | (stuck & ~restart)
    gn = np.linalg.norm(g, axis=1)
    return x, f, g, it, gn < gtol


_PROBE_DIRS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
                       + [[a, b, c] for a in (1, -1) for b in (1, -1) for c in (1, -1)], dtype=np.float64)
_PROBE_DIRS /= np.linalg.norm(_PROBE_DIRS, axis=1)[:, None]


def _minimise(obj, x0, gtol, max_iter, probe=0.05, rounds=4):
    """L-BFGS to a stationary point, then a second-order check: the energy
    at 14 points probe A away (axes and diagonals) must not be lower; a pose
    that sits on a saddle or maximum (a zero gradient by symmetry, say)
    resumes from its lowest probe.  Last, a shell search and a Nelder-Mead
    polish (the reference's own derivative-free method, src/energy_min.jl:52)
    from each end point replace it where lower (``_step_search``)."""
    x, f, g, it, conv = _lbfgs(obj, x0, max_iter=max_iter, gtol=gtol)
    total = it
    for _ in range(rounds):
        P = x.shape[0]
        trial = (x[:, None, :] + probe * _PROBE_DIRS[None, :, :]).reshape(-1, 3)
        ft, _ = obj(trial, np.repeat(np.arange(P), len(_PROBE_DIRS)))
        ft = ft.reshape(P, -1)
        ft = np.where(np.isfinite(ft), ft, np.inf)
        best = np.argmin(ft, axis=1)
        fb = ft[np.arange(P), best]
        with np.errstate(invalid="ignore"):
            lower = np.isfinite(f) & (fb < f - 1e-10 * (np.abs(f) + 1.0))
        if not lower.any():
            break
        x1 = x.copy()
        x1[lower] = trial.reshape(P, -1, 3)[lower, best[lower]]
        sub = np.nonzero(lower)[0]
        xs, fs, gs, its, cs = _lbfgs(lambda xx, ii: obj(xx, sub[ii]), x1[sub], max_iter=max_iter, gtol=gtol)
        x[sub], f[sub], g[sub], conv[sub] = xs, fs, gs, cs
        total += its
    # derivative-free polish across the cutoff's energy steps
    f_before = f.copy()
    x, f = _step_search(obj, x, f)
    moved = f < f_before
    if moved.any():
        mv = np.nonzero(moved)[0]
        fb, gb = obj(x[mv], mv)
        f[mv], g[mv] = fb, gb
    return x, f, g, total, conv


def _nelder_mead(obj, x0, f0, step=0.1, max_iter=1000, ftol=1e-10, xtol=1e-7):
    """Batched Nelder-Mead over P independent 3-dimensional problems, with
    the dimension-adaptive coefficients of Gao and Han (2012, Comput. Optim.
    Appl. 51:259) that Optim's NelderMead uses by default (reflection 1,
    expansion 1 + 2/n, contraction 3/4 - 1/(2n), shrink 1 - 1/n).  The
    initial simplex is x0 plus ``step`` A along each Cartesian axis.  A pose
    stops when the spread (standard deviation) of its vertex energies is
    below ``ftol`` K or every vertex lies within ``xtol`` A of its best.  At
    most three device calls per iteration serve all poses.  Returns the best
    vertex and its energy per pose."""
    n = 3
    a_r, a_e, a_c, a_s = 1.0, 1.0 + 2.0 / n, 0.75 - 0.5 / n, 1.0 - 1.0 / n
    x0 = np.asarray(x0, dtype=np.float64)
    P = x0.shape[0]
    S = np.repeat(x0[:, None, :], n + 1, axis=1)
    for i in range(n):
        S[:, i + 1, i] += step
    F = np.empty((P, n + 1))
    F[:, 0] = f0
    fv, _ = obj(S[:, 1:].reshape(-1, 3), np.repeat(np.arange(P), n))
    F[:, 1:] = fv.reshape(P, n)
    F = np.where(np.isfinite(F), F, np.inf)
    active = np.isfinite(F[:, 0])
    rows = np.arange(P)
    for _ in range(max_iter):
        o = np.argsort(F, axis=1, kind="stable")
        S = S[rows[:, None], o]
        F = F[rows[:, None], o]
        spread = np.where(np.isfinite(F).all(axis=1), np.std(F, axis=1), np.inf)
        size = np.max(np.linalg.norm(S[:, 1:] - S[:, :1], axis=2), axis=1)
        active &= ~((spread < ftol) | (size < xtol))
        if not active.any():
            break
        idx = np.nonzero(active)[0]
        c = S[idx, :n].mean(axis=1)
        worst = S[idx, n]
        xr = c + a_r * (c - worst)
        fr, _ = obj(xr, idx)
        fr = np.where(np.isfinite(fr), fr, np.inf)
        f0_, fn1, fn = F[idx, 0], F[idx, n - 1], F[idx, n]
        take_r = (fr >= f0_) & (fr < fn1)
        expand = fr < f0_
        outside = (fr >= fn1) & (fr < fn)
        inside = fr >= fn
        x2 = np.where(expand[:, None], c + a_e * (xr - c),
                      np.where(outside[:, None], c + a_c * (xr - c), c + a_c * (worst - c)))
        need = ~take_r
        f2 = np.full(len(idx), np.inf)
        if need.any():
            fv, _ = obj(x2[need], idx[need])
            f2[need] = np.where(np.isfinite(fv), fv, np.inf)
        new_x = np.where(take_r[:, None], xr, worst)
        new_f = np.where(take_r, fr, fn)
        # expansion: the better of the expanded and the reflected point
        e_ok = expand & (f2 < fr)
        new_x = np.where(e_ok[:, None], x2, np.where((expand & ~e_ok)[:, None], xr, new_x))
        new_f = np.where(e_ok, f2, np.where(expand & ~e_ok, fr, new_f))
        oc_ok = outside & (f2 <= fr)
        ic_ok = inside & (f2 < fn)
        new_x = np.where((oc_ok | ic_ok)[:, None], x2, new_x)
        new_f = np.where(oc_ok | ic_ok, f2, new_f)
        shrink = (outside & ~oc_ok) | (inside & ~ic_ok)
        S[idx, n] = new_x
        F[idx, n] = new_f
        if shrink.any():
            sh = idx[shrink]
            S[sh, 1:] = S[sh, :1] + a_s * (S[sh, 1:] - S[sh, :1])
            fv, _ = obj(S[sh, 1:].reshape(-1, 3), np.repeat(sh, n))
            F[sh, 1:] = np.where(np.isfinite(fv), fv, np.inf).reshape(-1, n)
    best = np.argmin(F, axis=1)
    return S[rows, best], F[rows, best]


def _shell_points(n_dirs=64, radii=(0.004, 0.008, 0.012, 0.016, 0.02, 0.025, 0.032, 0.04)):
    """Offsets (A) on spheres of the given radii, n_dirs Fibonacci-lattice
    directions each (the same set for every pose)."""
    k = np.arange(n_dirs) + 0.5
    z = 1.0 - 2.0 * k / n_dirs
    phi = math.pi * (1.0 + math.sqrt(5.0)) * k
    r = np.sqrt(1.0 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return np.concatenate([rad * d for rad in radii])


_SHELL = _shell_points()


def _step_search(obj, x, f, rounds=3, step=0.01):
    """Search the shells around each end point for lower energies and, from
    the lowest point found, polish by Nelder-Mead; repeated while it helps.
    The hard cutoff makes the energy jump by 4 eps ((sigma/r_c)^12 -
    (sigma/r_c)^6) wherever an atom crosses a cutoff sphere; next to a
    stationary point of the smooth part the lowest energy can lie just across
    such a step, in a slab a few thousandths of an A thick, where no gradient
    points."""
    P = x.shape[0]
    active = np.isfinite(f)
    for _ in range(rounds):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        m = len(_SHELL)
        trial = (x[idx, None, :] + _SHELL[None, :, :]).reshape(-1, 3)
        ft, _ = obj(trial, np.repeat(idx, m))
        ft = np.where(np.isfinite(ft), ft, np.inf).reshape(len(idx), m)
        best = np.argmin(ft, axis=1)
        fb = ft[np.arange(len(idx)), best]
        start_x = x[idx].copy()
        start_f = f[idx].copy()
        lower = fb < start_f
        start_x[lower] = trial.reshape(len(idx), m, 3)[lower, best[lower]]
        start_f[lower] = fb[lower]
        xn, fn = _nelder_mead(lambda xx, ii: obj(xx, idx[ii]), start_x, start_f, step=step)
        improved = fn < f[idx]
        if not improved.any():
            break
        x[idx[improved]], f[idx[improved]] = xn[improved], fn[improved]
        active = np.zeros(P, dtype=bool)
        active[idx[improved]] = True
    return x, f


def _start_template(molecule, box):
    """Cartesian template of ``molecule`` (given Cartesian, or fractional in ``box``)."""
    m = molecule if not molecule.frac else molecule.copy()
    if molecule.frac:
        m.atoms_x = box.f_to_c @ molecule.atoms_x
        m.charges_x = box.f_to_c @ molecule.charges_x
        m.com = box.f_to_c @ molecule.com
        m.frac = False
    return m


def find_energy_minima(crystal, molecule, ljff, starts_xf, ewald_precision=1e-6, precision="fp64",
                       gtol=1e-4, max_iter=500):
    """Minimise the energy of ``molecule`` (its orientation as given) from
    every centre-of-mass start in ``starts_xf`` ((P, 3) unit-cell fractional)
    at once.  Returns (xf_min (P, 3) unit-cell fractional in [0, 1),
    E_min (P,) kJ/mol, info dict)."""
    rep = replication_factors(crystal.box, ljff)
    xtal = replicate(crystal, rep)
    tmpl = _start_template(molecule, crystal.box)
    obj = _Objective(xtal, tmpl, ljff, ewald_precision, precision)
    starts = np.atleast_2d(np.asarray(starts_xf, dtype=np.float64))
    x0 = (starts / np.array(rep)[None, :]) @ xtal.box.f_to_c.T
    x, f, g, it, conv = _minimise(obj, x0, gtol, max_iter)
    xf_rep = x @ xtal.box.c_to_f.T
    xf = np.mod(np.mod(xf_rep, 1.0) * np.array(rep)[None, :], 1.0)
    info = {"iterations": it, "device_calls": obj.n_evals, "converged": conv,
            "grad_norm_K_per_A": np.linalg.norm(g, axis=1)}
    return xf, f * K_TO_KJ, info


def find_energy_minimum(crystal, molecule, ljff, ewald_precision=1e-6, precision="fp64", gtol=1e-4,
                        max_iter=500):
    """``src/energy_min.jl:19-58``: the molecule (Cartesian, or fractional in
    the unit cell) at a good initial position; returns (molecule at the
    minimum, fractional in the unit cell, and the minimum energy in kJ/mol)."""
    if needs_rotations(molecule):
        print("needs rotations. does not optimize over configurations, only over center of mass")
    com_xf = molecule.com if molecule.frac else crystal.box.c_to_f @ molecule.com
    xf, e, _ = find_energy_minima(crystal, molecule, ljff, com_xf[None, :], ewald_precision, precision, gtol,
                                  max_iter)
    out = Frac(molecule, crystal.box) if not molecule.frac else molecule.copy()
    translate_to(out, xf[0])
    return out, float(e[0])


def find_energy_minimum_gridsearch(crystal, molecule, ljff, resolution=1.0, precision="fp64"):
    """``src/energy_min.jl:75-85``: energy grid (kJ/mol) over the unit cell,
    the molecule moved to the grid point of lowest energy (first in Julia's
    column-major order among ties) and that energy."""
    from .grid import energy_grid
    grid = energy_grid(crystal, molecule, ljff, resolution=resolution, verbose=True, precision=precision)
    flat = grid.data.ravel(order="F")
    k = int(np.argmin(flat))
    idx = np.unravel_index(k, grid.n_pts, order="F")
    xf = np.array([i / (n - 1) if n > 1 else 0.0 for i, n in zip(idx, grid.n_pts)])
    out = Frac(molecule, crystal.box)
    translate_to(out, xf)
    return out, float(flat[k])
