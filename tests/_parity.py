"""Shared parity checks of device energies against the oracle's FP64 energies.

Two bounds, stated per energy E = E_vdw + E_es of a group (a Widom block, a
set of poses, a grid), TOL = 1e-5 (mixed) or 1e-10 (FP64), T = 298 K:

* every energy, weighted by its Boltzmann weight relative to the group's
  largest, w_i / w_max = e^(-(E_i - E_min)/T):
  (w_i / w_max) |E_gpu - E_oracle| <= TOL * (|E_vdw| + |E_es| + T).  An error
  dE moves a weight by dE/T (src/henry.jl:236-243), so this bounds what each
  insertion can do to K_H and <U>: the group's dominant (lowest) energy is
  held to TOL (|E| + T) -- 0.14 K at E = -13,712 K, 3e-3 K at E ~ 0 -- and an
  insertion e^-10 lighter to e^10 times that.  FP32 pair terms each carry
  ~1e-7 relative rounding, so an energy that cancels to ~0 K from terms of
  10^4 K carries ~1e-3 K of error whatever the kernel: relative to |E| + 1 K
  no FP32 pair arithmetic holds it to 1e-5, and relative to |E| + T it does;
* the condition-scale bound, |E_gpu - E_oracle| <= TOL * (S + 1 K), where S is
  the oracle's sum of absolute pair contributions behind the energy
  (sum |LJ_ij| + sum_s |q_s| (sum_j |erfc term_sj| + |Fourier_s|)), for every
  finite energy where S is computed: each term carries relative rounding TOL,
  so the error of a sum is bounded by TOL * S, not TOL * |E|.

Overlap (+Inf) flags must be identical everywhere.  The error of each group's
dominant energy relative to |E_vdw| + |E_es| + 1 K is reported beside them.
"""
from __future__ import annotations

import sys

import numpy as np

TOL = {"fp64": 1e-10, "mixed": 1e-5}
WINDOW = 40.0        # weights within e^-WINDOW of the group's largest
T_WEIGHT = 298.0     # K, temperature of the weight window
TINY_S = 1e-290      # block sums below this are weights at the edge of underflow


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def weighted_mask(total, groups=None, temperature=T_WEIGHT, window=WINDOW):
    """Finite energies within ``window`` kT of their group's lowest energy."""
    total = np.asarray(total, dtype=float)
    fin = np.isfinite(total)
    mask = np.zeros(total.shape, dtype=bool)
    if groups is None:
        groups = np.zeros(total.shape, dtype=np.int64)
    groups = np.asarray(groups)
    for gid in np.unique(groups):
        sel = (groups == gid) & fin
        if not sel.any():
            continue
        emin = total[sel].min()
        mask |= sel & (total <= emin + window * temperature)
    return mask


def weight_ratio(total, groups=None, temperature=T_WEIGHT):
    """e^(-(E_i - E_min)/T) per energy, E_min its group's lowest finite energy
    (0 for infinite energies); returns (ratios, index of each group's lowest)."""
    total = np.asarray(total, dtype=float)
    fin = np.isfinite(total)
    out = np.zeros(total.shape)
    lows = []
    if groups is None:
        groups = np.zeros(total.shape, dtype=np.int64)
    groups = np.asarray(groups)
    for gid in np.unique(groups):
        sel = np.flatnonzero((groups == gid) & fin)
        if len(sel) == 0:
            continue
        k = sel[np.argmin(total[sel])]
        lows.append(int(k))
        out[sel] = np.exp(-(total[sel] - total[k]) / temperature)
    return out, np.array(lows, dtype=np.int64)


def check_energies(got_vdw, got_es, want_vdw, want_es, tol, what, cond_scale=None, groups=None,
                   temperature=T_WEIGHT):
    """Asserts the bounds above.  Returns (largest error of a group's dominant
    energy relative to |E_vdw| + |E_es| + 1 K, reported only; largest weighted
    error relative to |E_vdw| + |E_es| + T; largest error relative to S + 1 K
    or None)."""
    gv, ge, wv, we = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=float))
                                           for a in (got_vdw, got_es, want_vdw, want_es)))
    got, want = gv + ge, wv + we
    assert np.array_equal(np.isinf(got), np.isinf(want)), f"{what}: overlap flags differ"
    fin = np.isfinite(want)
    assert np.all(np.isfinite(got[fin])), f"{what}: non-finite device energy where the oracle's is finite"
    if groups is not None:
        groups = np.broadcast_to(np.asarray(groups), want.shape)
    ratio, lows = weight_ratio(want, groups, temperature)
    mag = np.abs(wv) + np.abs(we)
    with np.errstate(invalid="ignore"):
        d = np.where(fin, np.abs(got - want), 0.0)
    rel_dom = float((d[lows] / (mag[lows] + 1.0)).max()) if len(lows) else 0.0
    err = ratio * d / (mag + temperature)
    rel_w = float(err.max()) if err.size else 0.0
    if rel_w > tol:
        i = int(np.argmax(err))
        raise AssertionError(f"{what}: weighted error {rel_w:.3e} > {tol:.0e} (E = {want[i]:.6f} K, device "
                             f"{got[i]:.6f} K, weight {ratio[i]:.3e} of its group's largest)")
    rel_s = None
    if cond_scale is not None and fin.any():
        s = np.broadcast_to(np.asarray(cond_scale, dtype=float), want.shape)
        err = d[fin] / (s[fin] + 1.0)
        rel_s = float(err.max())
        assert rel_s <= tol, f"{what}: error {rel_s:.3e} of the condition scale > {tol:.0e}"
    return rel_dom, rel_w, rel_s


def describe(res):
    """One line of the three errors check_energies returns."""
    rd, rw, rs = res
    out = f"dominant {rd:.2e} of |E|+1 K, weighted {rw:.2e} of |E|+T"
    return out + (f", all {rs:.2e} of S+1 K" if rs is not None else "")


def block_sums_from_energies(vdw, es, n_per_block, temperature):
    """(n_blocks, 3) of (S, W_vdw, W_es) in FP64, src/henry.jl:236-243: Inf
    energies weigh 0 and are left out of W."""
    vdw = np.asarray(vdw, dtype=float).reshape(-1, n_per_block)
    es = np.asarray(es, dtype=float).reshape(-1, n_per_block)
    tot = vdw + es
    with np.errstate(over="ignore", invalid="ignore"):
        w = np.where(np.isfinite(tot), np.exp(-tot / temperature), 0.0)
        wv = np.where(np.isfinite(tot), w * vdw, 0.0)
        we = np.where(np.isfinite(tot), w * es, 0.0)
    return np.stack([w.sum(axis=1), wv.sum(axis=1), we.sum(axis=1)], axis=1)


def repulsive_block_allowance(vdw, es, n_per_block, tol, temperature):
    """Per block, (S allowance, <U> allowance) from the per-insertion
    tolerance: an error dE_i = tol (|E_vdw,i| + |E_es,i| + T) moves weight i by
    dE_i / T and <U> = sum w E / sum w directly by dE_i, so

      S:   sum_i w_i expm1(dE_i / T) / S           (blocks whose every insertion
                                                      is repulsive; 0 elsewhere)
      <U>: sum_i w_i dE_i (1 + |E_i - <U>| / T) / (S |<U>|)   (every block).

    The S allowance covers small samples of dense frameworks (K_H of order
    e^(-E/T) with E ~ 10^5 K); the <U> allowance matters only where <U> is
    small beside the pair terms it cancels from (E = 51 K from 4330 K of LJ and
    -4279 K of Ewald, say), and is ~tol elsewhere."""
    vdw = np.asarray(vdw, dtype=float).reshape(-1, n_per_block)
    es = np.asarray(es, dtype=float).reshape(-1, n_per_block)
    tot = vdw + es
    s_out = np.zeros(tot.shape[0])
    u_out = np.zeros(tot.shape[0])
    with np.errstate(over="ignore", invalid="ignore"):
        for b in range(tot.shape[0]):
            fin = np.isfinite(tot[b])
            if not fin.any():
                continue
            E = tot[b][fin]
            w = np.exp(-(E - E.min()) / temperature)
            dE = tol * (np.abs(vdw[b][fin]) + np.abs(es[b][fin]) + temperature)
            u = float((w * E).sum() / w.sum())
            if u != 0.0:
                u_out[b] = float((w * dE * (1.0 + np.abs(E - u) / temperature)).sum() / (w.sum() * abs(u)))
            if E.min() > 0.0 and np.exp(-E.min() / temperature) > 0.0:
                arg = np.minimum(dE / temperature, 700.0)
                s_out[b] = float((w * np.expm1(arg)).sum() / w.sum())
    return s_out, u_out


def check_block_sums(got, want, rtol, what, s_allow=None, u_allow=None):
    """K_H (S) and <U>, <U_vdw>, <U_es> per block within rtol (S within
    max(rtol, s_allow[b]) and <U> within max(rtol, u_allow[b]) where
    allowances are given); a block whose weights all vanish has S = 0 on both
    sides (<U> is NaN there, as in the reference).  Returns the largest
    relative errors (S, <U>)."""
    got, want = np.asarray(got, dtype=float), np.asarray(want, dtype=float)
    worst_s = worst_u = 0.0
    for b in range(want.shape[0]):
        s_w, s_g = want[b, 0], got[b, 0]
        if s_w < TINY_S:
            # every weight is zero or subnormal (all E > ~2.1e5 K at 298 K)
            assert s_g < TINY_S, f"{what}: block {b}: oracle S = {s_w!r}, device {s_g!r}"
            continue
        es = abs(s_g - s_w) / s_w
        u_w = (want[b, 1] + want[b, 2]) / s_w
        u_g = (got[b, 1] + got[b, 2]) / s_g
        eu = abs(u_g - u_w) / max(abs(u_w), 1e-300)
        worst_s, worst_u = max(worst_s, es), max(worst_u, eu)
        lim = rtol if s_allow is None else max(rtol, float(s_allow[b]))
        assert es <= lim, f"{what}: block {b}: K_H relative error {es:.3e} > {lim:.1e}"
        ulim = rtol if u_allow is None else max(rtol, float(u_allow[b]))
        assert eu <= ulim, f"{what}: block {b}: <U> relative error {eu:.3e} > {ulim:.1e}"
        for k, name in ((1, "vdw"), (2, "es")):
            p_w, p_g = want[b, k] / s_w, got[b, k] / s_g
            scale = abs(u_w) + abs(p_w)
            assert abs(p_g - p_w) <= ulim * scale, f"{what}: block {b}: <U_{name}> {p_g!r} vs {p_w!r}"
    return worst_s, worst_u


def derivative_sum_bounds(vdw, es, d, cond, n_per_block, tol, temperature):
    """Bounds on the error of the per-block weighted derivative sums
    G_b = sum_i w_i dE_i/d theta (w_i = e^(-E_i/T)) and of
    g = d ln K_H / d theta = -sum_b G_b / (T sum_b S_b), from the oracle's
    per-insertion derivatives d (n, species, 3) and their summand conditions
    cond (sum of the magnitudes of the pair and Fourier summands behind each
    derivative, oracle/dual.c).  With each summand carrying relative rounding
    tol, a derivative errs by up to tol * cond_i, and an energy by up to
    dE_i = tol (|E_vdw| + |E_es| + T), which moves weight i by dE_i / T:

      |dG_b| <= tol * sum_i w_i (cond_i + |d_i| (|E_vdw| + |E_es| + T) / T)
      |dg|   <= (sum_b |dG_b| + |sum_b G_b| dS / S) / (T S),
                 dS / S = tol * sum_i w_i (|E_vdw| + |E_es| + T) / (T S).

    Derivatives of a neutral probe's energy by a framework charge cancel
    across its sites (sum_s q_s phi_t(x_s)), so cond_i can exceed |d_i| by
    10^3-10^4 there; elsewhere it is within a few |d_i|.  Returns
    (per-block bounds (blocks, species, 3), bound on g (species, 3))."""
    vdw = np.asarray(vdw, dtype=float)
    es = np.asarray(es, dtype=float)
    tot = vdw + es
    fin = np.isfinite(tot)
    with np.errstate(over="ignore", invalid="ignore"):
        w = np.where(fin, np.exp(-np.where(fin, tot, 0.0) / temperature), 0.0)
        mag = np.where(fin, np.abs(np.where(fin, vdw, 0.0)) + np.abs(es) + temperature, 0.0)
    ww = w[:, None, None]
    per = tol * ww * (np.abs(cond) + np.abs(d) * (mag / temperature)[:, None, None])
    nb = tot.shape[0] // n_per_block
    bb = per.reshape(nb, n_per_block, *d.shape[1:]).sum(axis=1)
    S = w.sum()
    G = (ww * d).sum(axis=0)
    ds = tol * float((w * mag).sum()) / (temperature * S)
    bg = (bb.sum(axis=0) + np.abs(G) * ds) / (temperature * S)
    return bb, bg
