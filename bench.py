"""Widom-insertion throughput on one MI355X: rigid 3-site CO2 with Ewald
electrostatics in a synthetic 600-atom triclinic framework (replicated 2x2x2
to 4800 atoms), 298 K, r_c 12.8 A, Ewald precision 1e-6.  One step = one
Widom run of 10^6 insertions (5 blocks x 200,000) through pm_widom, inputs
resident on the device (the insertion stream is generated on the device).

Prints one JSON line: insertions/s of the mixed-precision path (value), the
plain FP64 path beside it, the roofline of the energy kernel (FP32 VALU
bound; algorithmic flops per insertion from the per-pair constants below
and the exact expected in-cutoff pair count) and the CPU baseline: the
oracle's reference-form restatement (e^{ik.r} tables over every replicated
charge and replicated-box k-vector, as the reference's _phi) timed on a
bounded sample on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Widom insertions/s on 1 MI355X, mixed vs FP64 GPU vs CPU, at <=1e-4 rel K_H"

# algorithmic flops per unit of work
C_GEOM = 21   # per (site, replicated atom) candidate pair
C_LJ = 10     # per in-cutoff LJ pair
C_SR = 20     # per in-cutoff Ewald real-space pair
C_LJ_GRAD = 2  # gradient runs, per in-cutoff LJ pair: the sums of u^2 - u and u
C_SR_GRAD = 2  # gradient runs, per in-cutoff real-space pair: the sum of erfc(alpha r) / r
C_K = 14      # per (charge site, unit-cell k-vector)
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6


def in_cutoff(n_rep, omega_rep, r_cut):
    """Expected atoms within r_cut of a uniformly placed site:
    N_rep * (4/3) pi r_c^3 / Omega_rep (the cutoff sphere fits in the box)."""
    return n_rep * (4.0 / 3.0) * math.pi * r_cut ** 3 / omega_rep


def flops_all_pairs(n_rep, omega_rep, r_cut, n_lj, n_q, n_k, sr_cut):
    """Per insertion, the reference's all-pairs formulation: every site
    against every replicated atom."""
    n_sites = max(n_lj, n_q)  # CO2: LJ and charge sites coincide
    return (n_sites * n_rep * C_GEOM + n_lj * in_cutoff(n_rep, omega_rep, r_cut) * C_LJ
            + n_q * in_cutoff(n_rep, omega_rep, sr_cut) * C_SR + n_q * n_k * C_K)


def cpu_model():
    """Model name of this host's CPU (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pm, case, xr, seed, target_s):
    from oracle import oracle as orc
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    lj_off, q_off = pm.site_offsets(co2)
    o = orc.OracleSystem((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma),
                         x.xf.T, [ff.index(s) for s in x.species], x.atom_q, ff.epsilon_table,
                         ff.sigma2_table, math.sqrt(ff.r2_cutoff), lj_off.T,
                         [ff.index(s) for s in co2.atom_species], q_off.T, co2.charges_q,
                         ewald_precision=case.ewald_precision)
    threads = orc.lib().orc_num_threads()
    # probe the cost, then size the sample to ~target_s of CPU time
    n0 = max(threads, 4)
    t = time.perf_counter()
    o.widom_energies(seed, 0, n0, es_mode=1)
    dt = time.perf_counter() - t
    n = max(n0, int(n0 * target_s / max(dt, 1e-6)))
    t = time.perf_counter()
    o.widom_energies(seed, n0, n, es_mode=1)
    dt = time.perf_counter() - t
    ref_rate = n / dt
    # structure-factor CPU form, same cores (separates algorithm from device gains)
    t = time.perf_counter()
    m = max(threads * 8, 64)
    o.widom_energies(seed, 0, m, es_mode=0)
    sk_rate = m / (time.perf_counter() - t)
    return {
        "value": ref_rate, "unit": "insertions/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
        "sample": (f"{n} insertions of this workload (stream offset {n0}), reference form of the "
                   f"Fourier sum (e^(ik.r) tables over {xr.n_atoms} replicated charges x "
                   f"{o.wt.size} replicated-box k-vectors per site), OpenMP over insertions"),
        "sk_form_insertions_per_s": sk_rate,
    }


GCMC_WORKLOAD = ("xe_sbmof1: Xe in SBMOF-1 (tests/golden), UFF, r_c 14 A, replicated to 4320 atoms, 298 K, "
                 "0.5 bar (about 71 molecules per chain), after 1500 untimed cycles from an empty box")
FILL_CYCLES = 1500


def _gcmc_setup(pm):
    g = os.path.join(ROOT, "tests", "golden")
    x = pm.read_cif(os.path.join(g, "crystals", "SBMOF-1.cif"))
    ff = pm.LJForceField.from_csv(os.path.join(g, "forcefields", "UFF.csv"))
    xe = pm.Molecule.from_dir(os.path.join(g, "molecules", "Xe"))
    return x, ff, xe


def gcmc_leg(pm, cycles=100):
    """Grand-canonical chains through pm_gcmc_run: Markov steps per second
    with one chain per CU (256 chains) and with a single chain."""
    x, ff, xe = _gcmc_setup(pm)
    xr = pm.replicate(x, pm.replication_factors(x.box, ff))
    fw = pm.device.framework(xr, ff)
    out = {"workload": GCMC_WORKLOAD, "unit": "MC steps/s"}
    for key, n in (("chains_256", 256), ("chain_1", 1)):
        ch = pm.device.GcmcChains(fw, ff, [xe], [298.0] * n, [[0.5e5]] * n, 10 ** 6, 10 ** 6, capacity=512,
                                  seed=1)
        ch.run(FILL_CYCLES)
        t0 = np.array([ch.state(c)["markov_time"] for c in range(n)])
        t = time.perf_counter()
        ch.run(cycles)
        dt = time.perf_counter() - t
        per_chain = np.array([ch.state(c)["markov_time"] for c in range(n)]) - t0
        moves = int(per_chain.sum())
        mean_n = float(np.mean([ch.state(c)["n"][0] for c in range(n)]))
        # a launch lasts as long as its longest chain
        out[key] = {"chains": n, "cycles": cycles, "steps": moves, "seconds": dt, "value": moves / dt,
                    "us_per_step_longest_chain": dt / int(per_chain.max()) * 1e6, "mean_molecules": mean_n}
        print(f"gcmc {key}: {moves / dt:.3e} steps/s, {dt / int(per_chain.max()) * 1e6:.2f} us/step on the "
              f"longest chain, <N> {mean_n:.1f}", file=sys.stderr, flush=True)
    return out


def cpu_gcmc_baseline(pm, target_s):
    """The oracle's chain (oracle/gcmc.c) on the GCMC leg's workload, one
    independent chain per host thread (the oracle's OpenMP thread count; each
    chain's C loop runs with the GIL released): Markov steps per second summed
    over the chains, as the device figure sums over its chains."""
    import threading

    from oracle import gcmc as ogc
    from oracle import oracle as orc
    x, ff, xe = _gcmc_setup(pm)
    n_thr = orc.lib().orc_num_threads()

    def chain(seed):
        return ogc.OracleChain((x.box.a, x.box.b, x.box.c, x.box.alpha, x.box.beta, x.box.gamma), x.xf.T,
                               [ff.index(s) for s in x.species], ff.epsilon_table, ff.sigma2_table,
                               math.sqrt(ff.r2_cutoff), [((xe.atoms_x - xe.com[:, None]).T,
                                                          [ff.index(s) for s in xe.atom_species], False)],
                               298.0, [0.5e5], 10 ** 6, 10 ** 6, seed, capacity=512)

    chains = [chain(1 + c) for c in range(n_thr)]

    def fill(oc):
        oc.run(FILL_CYCLES)

    ths = [threading.Thread(target=fill, args=(oc,)) for oc in chains]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    t0 = [oc.state.t for oc in chains]
    stop = time.perf_counter() + target_s

    def go(oc):
        while time.perf_counter() < stop:
            oc.run(10)

    t = time.perf_counter()
    ths = [threading.Thread(target=go, args=(oc,)) for oc in chains]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t
    steps = sum(oc.state.t - a for oc, a in zip(chains, t0))
    return {"value": steps / dt, "unit": "MC steps/s", "cores": n_thr, "kind": "port", "cpu_model": cpu_model(),
            "sample": (f"{steps} Markov steps of {n_thr} independent chains (one per host thread, seeds 1..{n_thr}) "
                       f"of the GCMC leg's workload after {FILL_CYCLES} cycles each, "
                       f"<N> {np.mean([oc.state.n[0] for oc in chains]):.1f}")}


def screening_leg(pm, steps, n_per_block=20000):
    """256 synthetic frameworks x (CH4, CO2) x 10^5 insertions in one
    pm_widom_batch call per step (mixed precision)."""
    t0 = time.perf_counter()
    cases = pm.synthetic.screening_cases()
    runs, flop_in, n_ins = [], 0.0, 0
    for i, c in enumerate(cases):
        x, ff = c.crystal, c.ljff
        rep = pm.replication_factors(x.box, ff)
        xr = pm.replicate(x, rep)
        fw = pm.device.framework(xr, ff)
        ffh = pm.device.forcefield(ff)
        ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=c.ewald_precision, verbose=False)
        ew = pm.device.ewald(xr, fw, ep)
        r_lj, r_sr = math.sqrt(ff.r2_cutoff), ep.sr_cutoff_r
        for name in ("CH4", "CO2"):
            m = c.probes[name]
            runs.append((fw, ffh, ew if m.n_charges else None, pm.device.probe(m, ff), 298.0, 4242 + i, 5,
                         n_per_block))
            n = 5 * n_per_block
            n_ins += n
            flop_in += n * (m.n_atoms * in_cutoff(xr.n_atoms, xr.box.Omega, r_lj) * C_LJ
                            + (m.n_charges * in_cutoff(xr.n_atoms, xr.box.Omega, r_sr) * C_SR
                               if m.n_charges else 0.0))
    setup_s = time.perf_counter() - t0
    pm.device.widom_batch(runs, "mixed")  # warmup (image lists, scratch)
    # the cell cross-section for this batch (untimed, one call per setting; the
    # setting schedules work only): the process-wide choice was tuned on the
    # single CO2 run, whose point density differs from these frameworks'
    before = pm.device.get_tuning()
    trial = {}
    for cp in (256.0, 320.0, 384.0, 448.0, 512.0):
        pm.device.set_tuning(cp, before["chunk_r_max"])
        pm.device.widom_batch(runs, "mixed")
        trial[cp] = pm.device.last_kernel_stats()[0]
    best_cp = min(trial, key=trial.get)
    pm.device.set_tuning(best_cp, before["chunk_r_max"])
    pipe, pair, cand = [], [], []
    t = time.perf_counter()
    for _ in range(steps):
        pm.device.widom_batch(runs, "mixed")
        pipe.append(pm.device.last_kernel_stats()[0])
        pk, pc = pm.device.last_pair_kernel_stats()
        pair.append(pk)
        cand.append(pc)
    dt = (time.perf_counter() - t) / steps
    p_ms = float(np.mean(pair))
    flop = C_GEOM * float(np.mean(cand)) + flop_in
    t = time.perf_counter()
    pm.device.widom_batch(runs, "fp64")
    dt64 = time.perf_counter() - t
    pm.device.set_tuning(before["cell_points"], before["chunk_r_max"])
    print(f"screening: {len(runs)} runs, {n_ins} insertions, {dt * 1e3:.1f} ms/call, pair kernel "
          f"{p_ms:.1f} ms, setup {setup_s:.1f} s, cell points {best_cp:.0f}", file=sys.stderr, flush=True)
    return {
        "workload": (f"screening: {len(cases)} synthetic triclinic frameworks (N ~ logU[50,3000] capped at "
                     f"0.05/A^3, lengths U[8,40] A, angles U[70,110] deg, 5 charged species) x (CH4, CO2) x "
                     f"{5 * n_per_block} insertions, one pm_widom_batch call"),
        "runs": len(runs), "insertions": n_ins, "value": n_ins / dt, "unit": "insertions/s",
        "ms_per_call": dt * 1e3, "device_pipeline_ms": float(np.mean(pipe)), "pair_kernel_ms": p_ms,
        "pair_kernel_tflops": flop / (p_ms * 1e-3) / 1e12,
        "roofline": roofline("point_pair_kernel<float, 0>", flop, p_ms, PEAK_FP32_TFLOPS, "screening",
                             note="flop, time and traffic per call: the sum over the call's launch groups"),
        "fp64_value": n_ins / dt64, "fp64_calls": 1, "host_setup_s": setup_s,
        "tuning": {"cell_points": best_cp, "chunk_r_max": before["chunk_r_max"],
                   "source": "fastest device pipeline of one untimed call at each of 256-512 cell points",
                   "trial_ms": {str(int(k)): v for k, v in trial.items()}},
    }


def grids_leg(pm, steps):
    """Energy grids through pm_energy_grid (50^3
    CH4 grid of a 424-atom framework; 200^3 point-charge and CH4 grids of a
    3000-atom framework) and pm_energy_grid_batch (32 point-charge grids at
    80^3).  Grid points/s over the device pipeline (the grid is then copied
    to the caller's host buffer, which the wall-clock rates include: into a
    buffer reused across calls, and into a new one per call, whose fresh pages
    the operating system zeroes on first touch), and the pair kernel's
    algorithmic TFLOP/s with the per-pair constants above."""
    from porousmaterials_jl_amd.grid import _grid_setup

    def pair_flop(case, probe, cand, npts):
        x, ffl = case.crystal, case.ljff
        rep = pm.replication_factors(x.box, ffl)
        n_rep, omega = x.n_atoms * int(np.prod(rep)), x.box.Omega * float(np.prod(rep))
        r_c = math.sqrt(ffl.r2_cutoff)
        f = C_GEOM * cand
        if probe.n_atoms:
            f += npts * C_LJ * in_cutoff(n_rep, omega, r_c)
        if probe.n_charges and pm.has_charges(x):
            f += npts * C_SR * in_cutoff(n_rep, omega, r_c)
        return f

    def measure(key, case, probe_name, precision):
        probe = case.probes[probe_name]
        run, _ = _grid_setup(case.crystal, probe, case.ljff, case.meta["n_pts"], "K", False, False,
                             case.ewald_precision, electrostatics=True)
        buf = pm.device.energy_grid(*run, precision)  # warmup: image list, scratch; the reused host buffer
        pipe, pair, cand = [], [], []
        t = time.perf_counter()
        for _ in range(steps):
            pm.device.energy_grid(*run, precision, out=buf)
            pipe.append(pm.device.last_kernel_stats()[0])
            pk, pc = pm.device.last_pair_kernel_stats()
            pair.append(pk)
            cand.append(pc)
        wall = (time.perf_counter() - t) / steps
        t = time.perf_counter()
        for _ in range(steps):
            pm.device.energy_grid(*run, precision)
        wall_new = (time.perf_counter() - t) / steps
        npts = float(np.prod(case.meta["n_pts"]))
        p_ms = float(np.mean(pair))
        peak = PEAK_FP32_TFLOPS if precision == "mixed" else PEAK_FP64_TFLOPS
        kernel = "point_pair_kernel<float, 0>" if precision == "mixed" else "point_pair_kernel<double, 0>"
        rl = roofline(kernel, pair_flop(case, probe, float(np.mean(cand)), npts), p_ms, peak,
                      f"grid_{key}_{precision}")
        tf = rl["achieved"]
        return {"points": int(npts), "value": npts / (float(np.mean(pipe)) * 1e-3), "unit": "grid points/s",
                "wall_value": npts / wall, "wall_value_new_buffer": npts / wall_new,
                "device_pipeline_ms": float(np.mean(pipe)), "pair_kernel_ms": p_ms,
                "pair_kernel_tflops": tf, "frac": tf / peak, "candidates_per_point": float(np.mean(cand)) / npts,
                "roofline": rl}

    out = {}
    small = pm.synthetic.lj_grid_case()
    large = pm.synthetic.large_grid_case()
    for key, case, probe_name in (("ch4_50", small, "CH4"), ("ion_200", large, "ion"), ("ch4_200", large, "CH4")):
        out[key] = {p: measure(key, case, probe_name, p) for p in ("mixed", "fp64")}
        print(f"grid {key}: mixed {out[key]['mixed']['value']:.3e} points/s, fp64 "
              f"{out[key]['fp64']['value']:.3e} points/s", file=sys.stderr, flush=True)
    cases = pm.synthetic.small_grid_cases()
    runs = [_grid_setup(c.crystal, c.probes["ion"], c.ljff, c.meta["n_pts"], "K", False, False,
                        c.ewald_precision, electrostatics=True)[0] for c in cases]
    bufs = pm.device.energy_grid_batch(runs, "mixed")
    pipe = []
    t = time.perf_counter()
    for _ in range(steps):
        pm.device.energy_grid_batch(runs, "mixed", outs=bufs)
        pipe.append(pm.device.last_kernel_stats()[0])
    wall = (time.perf_counter() - t) / steps
    t = time.perf_counter()
    for _ in range(steps):
        pm.device.energy_grid_batch(runs, "mixed")
    wall_new = (time.perf_counter() - t) / steps
    npts = float(sum(np.prod(c.meta["n_pts"]) for c in cases))
    out["ion_80_x32_batch"] = {"points": int(npts), "value": npts / (float(np.mean(pipe)) * 1e-3),
                               "unit": "grid points/s", "wall_value": npts / wall,
                               "wall_value_new_buffer": npts / wall_new,
                               "device_pipeline_ms": float(np.mean(pipe)), "precision": "mixed"}
    out["workload"] = ("ch4_50: CH4 (LJ sphere) on a 50^3 grid of a cubic 25.8 A 424-atom framework (4 species); "
                       "ion_200 / ch4_200: unit point charge / CH4 on a 200^3 grid of a 3000-atom triclinic "
                       "framework (lengths 30-40 A, charged); ion_80_x32_batch: unit point charge on 80^3 grids "
                       "of 32 frameworks of 200-1500 atoms in one pm_energy_grid_batch call; r_c 12.8 A, Ewald "
                       "precision 1e-6")
    return out


def adjoint_leg(pm, case, fw, ff, ew, pr, xr, ep, n_per_block, seed, steps, minimise=True):
    """Gradients on the CO2 workload: (b) the adjoint Widom pass over the
    full 10^6 insertions (pm_widom_gradient, mixed; ``steps`` timed calls),
    (a) forces of 4096 CO2 poses (pm_pose_gradient, mixed) and, unless
    ``minimise`` is off, their batched minimisation (FP64)."""
    co2 = case.probes["CO2"]
    pm.device.widom_gradient(fw, ff, ew, pr, case.temperature, seed, 5, n_per_block, "mixed", case.ljff)
    pipe, pair, cands = [], [], []
    t = time.perf_counter()
    for _ in range(steps):
        sums, grads = pm.device.widom_gradient(fw, ff, ew, pr, case.temperature, seed, 5, n_per_block, "mixed",
                                               case.ljff)
        pipe.append(pm.device.last_kernel_stats()[0])
        pk, pc = pm.device.last_pair_kernel_stats()
        pair.append(pk)
        cands.append(pc)
    dt_b = (time.perf_counter() - t) / steps
    k_ms, p_ms, cand = float(np.mean(pipe)), float(np.mean(pair)), float(np.mean(cands))
    n_ins = 5 * n_per_block
    r_lj, r_sr = math.sqrt(case.ljff.r2_cutoff), ep.sr_cutoff_r
    flop = C_GEOM * cand + n_ins * (co2.n_atoms * in_cutoff(xr.n_atoms, xr.box.Omega, r_lj) * (C_LJ + C_LJ_GRAD)
                                    + co2.n_charges * in_cutoff(xr.n_atoms, xr.box.Omega, r_sr)
                                    * (C_SR + C_SR_GRAD))
    nt = len({case.ljff.index(sp) for sp in case.crystal.species})
    nt_kernel = 2 if nt <= 2 else 4 if nt <= 4 else 6 if nt <= 6 else 8
    S = float(sums[:, 0].sum())
    g = -grads.sum(axis=0) / (case.temperature * S)
    lj_off, _ = pm.site_offsets(co2)
    rng = np.random.Generator(np.random.PCG64(4096))
    com = rng.random((4096, 3)) @ xr.box.f_to_c.T
    poses = []
    for c in com:
        R = pm.arvo_rotation_matrix(*rng.random(3))
        poses.append((c[None, :] + lj_off.T @ R.T) @ xr.box.c_to_f.T)
    poses = np.array(poses)
    species = [case.ljff.index(s) for s in co2.atom_species]
    pm.device.pose_gradients(fw, ff, ew, species, list(co2.charges_q), poses, "mixed")
    t = time.perf_counter()
    pm.device.pose_gradients(fw, ff, ew, species, list(co2.charges_q), poses, "mixed")
    dt_a = time.perf_counter() - t
    a_ms = pm.device.last_kernel_stats()[0]
    out = {
        "widom_gradient": {"insertions": n_ins, "value": n_ins / dt_b, "steps": steps,
                           "unit": "insertions/s", "device_pipeline_ms": k_ms, "pair_kernel_ms": p_ms,
                           "parameters": f"d ln K_H / d(eps_t, sigma_t, q_t), {nt} framework species",
                           "dlnKH_deps_T1": float(g[case.ljff.index("T1"), 0]),
                           "roofline": roofline(f"point_pair_kernel<float, {nt_kernel}>", flop, p_ms,
                                                PEAK_FP32_TFLOPS, "adjoint")},
        "pose_forces": {"poses": 4096, "value": 4096 / dt_a, "unit": "poses/s", "device_ms": a_ms},
    }
    msg = f"adjoint: pass {dt_b * 1e3:.2f} ms (device {k_ms:.2f} ms, pair kernel {p_ms:.2f} ms), " \
          f"4096 pose forces {dt_a * 1e3:.1f} ms"
    if minimise:
        rep = np.array(pm.replication_factors(case.crystal.box, case.ljff))
        starts = np.mod((com @ xr.box.c_to_f.T) * rep[None, :], 1.0)
        t = time.perf_counter()
        _, e_min, info = pm.find_energy_minima(case.crystal, co2, case.ljff, starts, max_iter=200)
        dt_m = time.perf_counter() - t
        out["minimise_4096"] = {"seconds": dt_m, "iterations": int(info["iterations"]),
                                "device_calls": int(info["device_calls"]),
                                "converged": int(np.sum(info["converged"]))}
        msg += f", 4096 minimisations {dt_m:.1f} s"
    print(msg, file=sys.stderr, flush=True)
    return out


def fp64_leg(pm, run, n_ins, nb, case, sums32, value32, steps, flop_of):
    """The headline workload through the plain FP64 path: insertions/s over
    ``steps`` calls, the pair kernel's algorithmic TFLOP/s against the FP64
    peak, and K_H / <U> of the mixed run (``sums32``; one untimed mixed call
    when None) relative to the FP64 run."""
    run("fp64")
    k64s, p64s, c64s = [], [], []
    t = time.perf_counter()
    for _ in range(steps):
        s64 = run("fp64")
        k64s.append(pm.device.last_kernel_stats()[0])
        pk, pc = pm.device.last_pair_kernel_stats()
        p64s.append(pk)
        c64s.append(pc)
    dt64 = (time.perf_counter() - t) / steps
    k64, p64 = float(np.mean(k64s)), float(np.mean(p64s))
    if sums32 is None:
        sums32 = run("mixed")
    r64 = pm.block_statistics(s64, nb, case.temperature, 1.0)
    r32 = pm.block_statistics(sums32, nb, case.temperature, 1.0)
    kh = "henry coefficient [mol/(m³-bar)]"
    rl = roofline("point_pair_kernel<double, 0>", flop_of(float(np.mean(c64s))), p64, PEAK_FP64_TFLOPS,
                  "widom_fp64")
    tf = rl["achieved"]
    out = {"value": n_ins / dt64, "steps": steps, "device_pipeline_ms": k64, "pair_kernel_ms": p64,
           "pair_kernel_tflops": tf, "peak_tflops": PEAK_FP64_TFLOPS, "frac": tf / PEAK_FP64_TFLOPS,
           "roofline": rl,
           "kh_rel_diff_mixed_vs_fp64": abs(r32[kh] - r64[kh]) / abs(r64[kh]),
           "u_rel_diff_mixed_vs_fp64": abs(r32["⟨U⟩ (K)"] - r64["⟨U⟩ (K)"]) / abs(r64["⟨U⟩ (K)"])}
    if value32 is not None:
        out["mixed_over_fp64"] = value32 / (n_ins / dt64)
    print(f"fp64: {steps} steps, {dt64 * 1e3:.3f} ms/step, device pipeline {k64:.3f} ms, pair kernel "
          f"{p64:.3f} ms", file=sys.stderr, flush=True)
    return out


def _traffic_record(key):
    """HBM bytes per call of a leg's pair kernel, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (scaled by a plain copy of known size)
    and committed in profiles/kernel_traffic.json under ``key``; None if
    absent."""
    path = os.path.join(ROOT, "profiles", "kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def roofline(kernel, flop, kernel_ms, peak, traffic_key, **extra):
    """The roofline object of a leg's dominant kernel: algorithmic TFLOP/s
    over the kernel's HIP-event time against the published VALU peak, and
    its measured HBM bytes per call when committed."""
    achieved = flop / (kernel_ms * 1e-3) / 1e12
    out = {"bound": "valu-fp32" if peak == PEAK_FP32_TFLOPS else "valu-fp64", "achieved": achieved, "peak": peak,
           "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None, "kernel": kernel, "kernel_ms": kernel_ms,
           "flop_per_launch": flop, **extra}
    t = _traffic_record(traffic_key)
    if t is not None:
        out["traffic"] = t["bytes_per_launch"]
        out["traffic_source"] = t["source"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-per-block", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-screening", action="store_true", help="skip the batched screening leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (gradient) leg")
    ap.add_argument("--no-grids", action="store_true", help="skip the energy-grid leg")
    ap.add_argument("--no-gcmc", action="store_true", help="skip the grand-canonical Monte Carlo leg")
    ap.add_argument("--no-minimise", action="store_true",
                    help="adjoint leg: skip the batched minimisation of 4096 poses")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (pm_copy_probe) three times: the "
                         "calibration launch for FETCH_SIZE / WRITE_SIZE counter runs")
    ap.add_argument("--cell-points", type=float, default=None,
                    help="chunk tuning: points per Widom cell (default: autotuned on the workload)")
    ap.add_argument("--chunk-r-max", type=float, default=None,
                    help="chunk tuning: largest chunk radius in A (default: autotuned on the workload)")
    ap.add_argument("--only", choices=("fp64", "adjoint", "screening", "grids", "gcmc"), default=None,
                    help="run one side leg alone (no headline timing, no autotuning: the library's default "
                         "chunk shape unless given) and print its JSON object; for per-leg kernel traces")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(Widom blocks would shard by insertion range with one 120-byte reduction).",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    pm = __graft_entry__.load_package()
    case = pm.synthetic.co2_widom_case(n_per_block=args.n_per_block)
    x, ff, co2 = case.crystal, case.ljff, case.probes["CO2"]
    rep = pm.replication_factors(x.box, ff)
    xr = pm.replicate(x, rep)
    ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision)
    fw = pm.device.framework(xr, ff)
    ew = pm.device.ewald(xr, fw, ep)
    pr = pm.device.probe(co2, ff)
    n_k = pm.device.ewald_info(ew)
    nb = args.n_per_block
    seed = case.meta["insertion_seed"]
    n_ins = 5 * nb

    def run(precision):
        return pm.device.widom(fw, ff, ew, pr, case.temperature, seed, 5, nb, precision)

    if args.only is not None:
        if args.cell_points is not None or args.chunk_r_max is not None:
            pm.device.set_tuning(args.cell_points, args.chunk_r_max)
        r_lj, r_sr = math.sqrt(ff.r2_cutoff), ep.sr_cutoff_r
        n_pts = n_ins * max(co2.n_atoms, co2.n_charges)
        if args.only == "fp64":
            res = fp64_leg(pm, run, n_ins, nb, case, None, None, max(1, min(args.steps, 5)),
                           lambda c: C_GEOM * c + n_ins * (co2.n_atoms * in_cutoff(xr.n_atoms, xr.box.Omega, r_lj)
                                                           * C_LJ + co2.n_charges
                                                           * in_cutoff(xr.n_atoms, xr.box.Omega, r_sr) * C_SR))
        elif args.only == "adjoint":
            res = adjoint_leg(pm, case, fw, ff, ew, pr, xr, ep, nb, seed, max(1, min(args.steps, 5)),
                              minimise=not args.no_minimise)
        elif args.only == "screening":
            res = screening_leg(pm, max(1, min(args.steps, 3)))
        elif args.only == "grids":
            res = grids_leg(pm, max(1, min(args.steps, 3)))
        else:
            res = gcmc_leg(pm)
        print(json.dumps({"leg": args.only, "points": n_pts, **res}), flush=True)
        return

    # chunk shapes: autotuned on this workload (untimed) unless given; they
    # change the schedule only, never a result
    if args.cell_points is None and args.chunk_r_max is None:
        tuning = pm.device.autotune_widom(fw, ff, ew, pr, case.temperature, seed, 5, nb, "mixed")
        tuning["source"] = "pm_autotune_widom on this workload"
    else:
        pm.device.set_tuning(args.cell_points, args.chunk_r_max)
        tuning = dict(pm.device.get_tuning(), source="command line")
    print(f"tuning: {tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        run("mixed")
    pipe_ms, pair_ms, cands = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sums = run("mixed")
        pipe_ms.append(pm.device.last_kernel_stats()[0])
        pk, pc = pm.device.last_pair_kernel_stats()
        pair_ms.append(pk)
        cands.append(pc)
    elapsed = time.perf_counter() - t0
    value = args.steps * n_ins / elapsed
    k_ms = float(np.mean(pipe_ms))
    p_ms = float(np.mean(pair_ms))
    cand = float(np.mean(cands))
    print(f"mixed: {args.steps} steps, {elapsed / args.steps * 1e3:.3f} ms/step, device pipeline "
          f"{k_ms:.3f} ms, pair kernel {p_ms:.3f} ms", file=sys.stderr, flush=True)

    r_lj, r_sr = math.sqrt(ff.r2_cutoff), ep.sr_cutoff_r
    n_points = n_ins * max(co2.n_atoms, co2.n_charges)
    # pair kernel: C_GEOM per evaluated (point, candidate image) pair, C_LJ and
    # C_SR per in-cutoff pair
    pair_flop = (C_GEOM * cand + n_ins * (co2.n_atoms * in_cutoff(xr.n_atoms, xr.box.Omega, r_lj) * C_LJ
                                         + co2.n_charges * in_cutoff(xr.n_atoms, xr.box.Omega, r_sr) * C_SR))
    flop_ap = flops_all_pairs(xr.n_atoms, xr.box.Omega, r_lj, co2.n_atoms, co2.n_charges, n_k, r_sr)
    out = {
        "metric": METRIC, "value": value, "unit": "insertions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 wrap/r^2/Fourier/Boltzmann sums", "data": "synthetic",
        "config": {
            "workload": ("co2_widom: rigid 3-site CO2 at 298 K, Ewald precision 1e-6, r_c 12.8 A, "
                         f"synthetic triclinic 600-atom framework replicated {rep} to "
                         f"{xr.n_atoms} atoms, {n_k} unit-cell k-vectors, "
                         f"{n_ins} insertions = 5 blocks x {nb}"),
            "insertions_per_step": n_ins, "precision": "mixed",
            "tuning": {k: tuning[k] for k in ("cell_points", "chunk_r_max", "source")}},
        "roofline": roofline("point_pair_kernel<float, 0>", pair_flop, p_ms, PEAK_FP32_TFLOPS, "widom_mixed",
                             candidate_pairs_per_point=cand / n_points,
                             pruning_factor=(n_points * xr.n_atoms) / cand),
        "device_pipeline_ms": k_ms,
        "all_pairs_equivalent_tflops": flop_ap * n_ins / (k_ms * 1e-3) / 1e12,
    }
    if not args.no_fp64:
        out["fp64"] = fp64_leg(pm, run, n_ins, nb, case, sums, value, max(1, min(args.steps, 5)),
                               lambda c: pair_flop)
        out["mixed_over_fp64"] = out["fp64"].pop("mixed_over_fp64")
        out["kh_rel_diff_mixed_vs_fp64"] = out["fp64"].pop("kh_rel_diff_mixed_vs_fp64")
        out["u_rel_diff_mixed_vs_fp64"] = out["fp64"].pop("u_rel_diff_mixed_vs_fp64")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(pm, case, fw, ff, ew, pr, xr, ep, nb, seed, max(1, min(args.steps, 5)),
                                     minimise=not args.no_minimise)
    if not args.no_screening:
        out["screening"] = screening_leg(pm, max(1, min(args.steps, 3)))
    if not args.no_grids:
        out["grids"] = grids_leg(pm, max(1, min(args.steps, 3)))
    if not args.no_gcmc:
        out["gcmc"] = gcmc_leg(pm)
    if not args.no_cpu:
        cb = cpu_baseline(pm, case, xr, seed, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
        if not args.no_gcmc:
            cb["gcmc"] = cpu_gcmc_baseline(pm, min(5.0, args.cpu_seconds))
            out["gcmc"]["gpu_over_cpu"] = out["gcmc"]["chains_256"]["value"] / cb["gcmc"]["value"]
    if args.copy_probe:
        ms = ctypes.c_double()
        _lib = pm._lib
        _lib.check(_lib.load().pm_copy_probe(256 << 20, 3, ctypes.byref(ms)))
        out["copy_probe"] = {"bytes": 256 << 20, "ms": ms.value,
                             "GB_per_s": 2 * (256 << 20) / (ms.value * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
