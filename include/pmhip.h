/*
 * pmhip.h -- C ABI of the MI355X (gfx950) guest-host energy library
 * libpmhip.so: Lennard-Jones + Ewald energies of rigid adsorbates in
 * periodic frameworks, Widom-insertion block sums and energy grids.
 *
 * The reference (PorousMaterials.jl v0.4.2) is pure Julia and has no FFI of
 * its own; each entry point below names the Julia function whose numerical
 * work it replaces, so a Julia `ccall` shim can keep the reference's call
 * surface.  Paths are relative to the reference tree.
 *
 * Conventions
 *   - Every function returns PM_OK (0) or a PM_ERR_* code; the message of
 *     the last failure on the calling thread is pm_last_error().
 *   - 3x3 matrices are column-major (Julia layout): m[row + 3*col].
 *     f_to_c must be upper-triangular (Xtals Box convention, a along x).
 *   - Coordinates are AoS xyz per point (a Julia 3xN Matrix{Float64}).
 *   - Caller buffers are borrowed for the duration of a call only; device
 *     state lives in the opaque handles until the matching *_destroy.
 *   - Calls are synchronous.  All calls are serialised by one library
 *     mutex, so any thread may call; handles must not be destroyed while in
 *     use by another thread.
 *   - precision: PM_FP64 computes every pair in FP64; PM_MIXED computes pair
 *     arithmetic (LJ, erfc/r) in FP32 with FP64 fractional wrap, r^2 and
 *     cutoff tests, FP64 per-point accumulation, FP64 Fourier phases and
 *     structure factors and FP64 Boltzmann sums.
 */
#ifndef PMHIP_H
#define PMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PM_OK = 0,
    PM_ERR_ARG = 1,     /* invalid argument (shape, index, range)      */
    PM_ERR_DEVICE = 2,  /* HIP runtime error                           */
    PM_ERR_OOM = 3,     /* device allocation failed                    */
    PM_ERR_STATE = 4    /* library not initialised / bad handle        */
};

enum { PM_FP64 = 0, PM_MIXED = 1 };

#define PM_MAX_SITES 8          /* sites (LJ spheres + point charges) per probe */
#define PM_MAX_FW_SPECIES 64    /* distinct LJ species within one framework     */
#define PM_MAX_GRAD_TYPES 8     /* framework species of a pm_widom_gradient run  */

typedef struct pm_framework pm_framework;
typedef struct pm_forcefield pm_forcefield;
typedef struct pm_ewald pm_ewald;
typedef struct pm_probe pm_probe;
typedef struct pm_access pm_access;

/* Library version string. */
const char* pm_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* pm_last_error(void);

/* Select the HIP device and create the library stream.  Must precede every
 * other call except pm_version/pm_last_error.  Idempotent per device. */
int pm_init(int32_t device);

/*
 * Framework = the crystal replicated by `rep` (src/henry.jl:80-86,
 * src/grid.jl:331-332, Xtals `replicate`).  f_to_c is the UNIT cell's;
 * xf (n_atoms x 3) are unit-cell fractional coordinates; species are LJ
 * species indices into the force field's tables; charge may be NULL
 * (uncharged framework).  The library builds the replicated atom list in
 * Xtals order (ra outermost, then rb, rc, then atoms).
 * Replaces the data preparation of `_conduct_Widom_insertions`
 * (src/henry.jl:185-203) and `energy_grid` (src/grid.jl:326-332).
 */
int pm_framework_create(const double f_to_c[9], const int32_t rep[3], int64_t n_atoms,
                        const double* xf, const int32_t* species, const double* charge,
                        pm_framework** out);
int pm_framework_destroy(pm_framework* fw);
/* Replicated atom count, number of distinct species, total charge. */
int pm_framework_info(const pm_framework* fw, int64_t* n_replicated, int32_t* n_species,
                      double* net_charge);

/*
 * LJ force field as dense pair tables over its species (the Dict-of-Dict
 * `LJForceField.epsilon` / `sigma^2` after mixing, src/forcefield.jl:12-22,
 * 46-100): epsilon[i*n + j] (K), sigma2[i*n + j] (A^2), r2_cutoff (A^2).
 */
int pm_forcefield_create(int32_t n_species, const double* epsilon, const double* sigma2,
                         double r2_cutoff, pm_forcefield** out);
int pm_forcefield_destroy(pm_forcefield* ff);

/*
 * Ewald set-up on the REPLICATED box (src/electrostatics.jl:256-272 output):
 * alpha (1/A), real-space cutoff (A) and the k-vector list exactly as
 * `precompute_kvec_wts` enumerates it (kidx n_k x 3 integer multiples of the
 * replicated box's reciprocal vectors, weights wt).  The library keeps the
 * vectors whose unit-cell structure factor does not vanish identically
 * (n divisible by rep) and computes S(k) on the device in FP64.
 * Replaces `setup_Ewald_sum` + `Eikr` (src/electrostatics.jl:90-101,256-272)
 * as the input to the Fourier sum.
 */
int pm_ewald_create(const pm_framework* fw, double alpha, double sr_cutoff, int64_t n_kvecs,
                    const int32_t* kidx, const double* wt, pm_ewald** out);
int pm_ewald_destroy(pm_ewald* ew);
/* Number of k-vectors kept after folding onto the unit cell. */
int pm_ewald_info(const pm_ewald* ew, int64_t* n_kept);

/*
 * Rigid probe (adsorbate template, src/molecule.jl:10-15,56-91): LJ spheres
 * (species index, Cartesian offset from the centre of mass) and point charges
 * (charge, Cartesian offset).  Coincident LJ and charge sites are merged.
 * n_lj + n_q (after merging) <= PM_MAX_SITES.
 */
int pm_probe_create(int32_t n_lj, const double* lj_xyz, const int32_t* lj_species,
                    int32_t n_q, const double* q_xyz, const double* q, pm_probe** out);
int pm_probe_destroy(pm_probe* pr);

/*
 * Accessibility grid (`Grid{Bool}` of `compute_accessibility_grid`,
 * src/grid.jl:12-18,689-774) over the UNIT cell: n_pts[0]*n_pts[1]*n_pts[2]
 * bytes, i fastest, nonzero = accessible.  A Widom run given one counts an
 * insertion whose centre of mass is not `accessible(grid, xf, repfactors)`
 * (src/grid.jl:778-836) as E_vdw = E_es = +Inf without computing its energy
 * (src/henry.jl:225-232).  1 <= n_pts[d] <= 65536, at most 2^31 points.
 */
int pm_access_create(const int32_t n_pts[3], const uint8_t* data, pm_access** out);
int pm_access_destroy(pm_access* acc);

/*
 * Pocket blocking of an energy grid: the steps of `compute_accessibility_grid`
 * after `energy_grid` (src/grid.jl:704-764), on the host.
 * data: n_pts[0]*n_pts[1]*n_pts[2] energies, i fastest, in the units of
 * energy_tol.  block_pockets = 0: out_accessible = data < energy_tol
 * (src/grid.jl:704-710).  block_pockets = 1: flood fill into segments
 * (`_flood_fill!`, `_segment_grid`, src/grid.jl:383-456: a seed needs data <= energy_tol, a
 * neighbour data < energy_tol, 26 neighbours, no periodic wrap), segment
 * connections across the three unit-cell faces (`_note_connection!`,
 * `_build_list_of_connections`, :485-530), every simple cycle of the segment / connection graph
 * (`_translate_into_graph`, `_classify_segments`, :533-629; Johnson's
 * algorithm), segments on a cycle with a nonzero net unit-cell translation
 * accessible, all other segments blocked (:632-646, 750-751).
 * out_accessible: 1 / 0 per point.  out_segments (nullable): the flood fill's
 * labels per point (1..n_segments, -1 not occupiable), before blocking.
 * out_counts (nullable): n_segments, segments blocked, simple cycles found.
 * Fails with PM_ERR_ARG past max_cycles simple cycles (0: 10^8).
 */
int pm_block_pockets(const int32_t n_pts[3], const double* data, double energy_tol, int32_t block_pockets,
                     int64_t max_cycles, uint8_t* out_accessible, int32_t* out_segments, int64_t out_counts[3]);

/*
 * Guest-host LJ energy (K) of n_poses rigid poses: `vdw_energy(crystal,
 * molecule, ljff)` (src/vdw_energetics.jl:25-47,62-64) per pose.
 * site_xf: n_poses x n_sites x 3, replicated-box fractional coordinates.
 * Overlap (r^2 < 0.1 A^2 inside the cutoff) gives +Inf for that pose.
 */
int pm_vdw_energy(const pm_framework* fw, const pm_forcefield* ff, int32_t n_sites,
                  const int32_t* site_species, int64_t n_poses, const double* site_xf,
                  int32_t precision, double* out);

/*
 * Guest-host Ewald energy (K) of n_poses sets of point charges:
 * `electrostatic_potential_energy(crystal, molecule, eparams, eikr)`
 * (src/electrostatics.jl:331-388) per pose, split into the real-space (sr)
 * and Fourier (lr) parts of the returned `EwaldSum`.  Summing over poses
 * gives `total_electrostatic_potential_energy` (src/electrostatics.jl:558-568).
 * charge_xf: n_poses x n_charges x 3, replicated-box fractional coordinates.
 */
int pm_electrostatic_energy(const pm_framework* fw, const pm_ewald* ew, int32_t n_charges,
                            const double* charges, int64_t n_poses, const double* charge_xf,
                            int32_t precision, double* out_sr, double* out_lr);

/*
 * Energies and forces of n_poses rigid poses: per pose out_e = (vdw, sr, lr)
 * as pm_vdw_energy / pm_electrostatic_energy give them, and out_grad
 * (n_poses x n_sites x 3) = dE/d(site Cartesian xyz) in K/A of the pose's
 * total energy E = vdw + sr + lr.  Sites carry an LJ species (-1: none) and a
 * charge (0: none); site_xf as for pm_vdw_energy.  Replaces the derivative-free
 * objective of `find_energy_minimum` (src/energy_min.jl:19-58, Nelder-Mead on
 * `vdw_energy`), so a gradient-based minimiser can use LJ + Ewald forces.
 * Gradients are undefined for poses whose energy is +Inf (overlap).
 */
int pm_pose_gradient(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, int32_t n_sites,
                     const int32_t* site_species, const double* site_charges, int64_t n_poses,
                     const double* site_xf, int32_t precision, double* out_e, double* out_grad);

/*
 * Widom insertions (src/henry.jl:185-265): n_blocks x n_per_block insertions
 * of the probe, insertion g = b*n_per_block + i drawing its centre of mass
 * (uniform in the replicated box) and an absolute Arvo rotation from the
 * Philox4x32-10 stream keyed by `seed`.  ew may be NULL (no electrostatics,
 * `charged_system == false`).  out_block_sums[3*b + {0,1,2}] =
 * (sum e^{-E/T}, sum E_vdw e^{-E/T}, sum E_es e^{-E/T}) with Inf energies
 * excluded from the weighted sums (src/henry.jl:236-243).
 * out_energies (nullable): n_blocks*n_per_block x 2 (E_vdw, E_es) in K.
 * access (nullable): the `accessibility_grid` argument of henry_coefficient
 * (src/henry.jl:38,225), over the unit cell of the framework's replication.
 */
int pm_widom(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
             const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
             int64_t n_per_block, int32_t precision, double* out_block_sums,
             double* out_energies, const pm_access* access);

/*
 * Energy grid (src/grid.jl:291-362, single-site probe): point (i,j,k) of an
 * n_pts grid including fractional 0 and 1 of the UNIT cell, energy in K of
 * the probe centred there; out is i-fastest (Julia Array{Float64,3}).
 * A charged single-site probe adds q * phi (the point-charge potential
 * energy of `electrostatic_potential_energy(crystal, ion(q, xf), ...)`),
 * which needs ew.  PM_MIXED keeps the device grid in FP32.
 */
int pm_energy_grid(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                   const pm_probe* pr, const int32_t n_pts[3], int32_t precision,
                   double* out);

/*
 * Adjoint of pm_widom: the same run (out_block_sums bit for bit as pm_widom
 * gives them) plus, per block b and force-field species t present in the
 * framework, the Boltzmann-weighted derivative sums
 *   out_block_grads[(b*n_ff + t)*3 + {0,1,2}] = sum_i w_i dE_i/d{eps_t, sigma_t, q_t}
 * with w_i = e^{-E_i/T} (Inf energies excluded), n_ff = the force field's species
 * count, zeros for species absent from the framework.  eps_t, sigma_t are the
 * framework-side pure parameters (the probe's held fixed, also where a probe
 * site shares the species); q_t is a uniform shift of the charges of every
 * framework atom of species t.  jac[(i*n_ff + j)*4 + k] is the mixing rule's
 * Jacobian for framework species i and probe species j: d eps_ij/d eps_i,
 * d eps_ij/d sigma_i, d sigma_ij/d eps_i, d sigma_ij/d sigma_i.  Then
 * d ln K_H / d theta = -(1/T) sum_b G_b(theta) / sum_b S_b (src/henry.jl:112-135
 * differentiated).  The reference has no derivative of henry_coefficient;
 * this is the backward pass of a ChainRulesCore rrule around it.
 * At most PM_MAX_GRAD_TYPES framework species.  access as for pm_widom
 * (blocked insertions carry zero weight and zero derivative).
 */
int pm_widom_gradient(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                      const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
                      int64_t n_per_block, int32_t precision, const double* jac,
                      double* out_block_sums, double* out_block_grads, const pm_access* access);

/*
 * pm_widom without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as pm_widom (block sums bit for bit), without
 * per-insertion energies.  Its buffers (point slots, cell lists, per-workgroup
 * partials, block sums, staged tables) live in a fixed device array of
 * PM_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
 * process ends; a run whose buffers exceed it fails with PM_ERR_ARG before any
 * device work (about 235 bytes per insertion of a 3-site probe: some 2.8e5 CO2
 * insertions per call).  The framework's image list for the force field's
 * cutoff must already exist (any pm_widom call with the same framework and
 * force field builds it); otherwise PM_ERR_STATE.  It calls no hipMalloc,
 * hipFree or allocating routine.  Calls are serialised by the library mutex
 * and complete before they return, so two calls never use the array at once;
 * it keeps the last call's contents, and the call clears what must start at
 * zero.
 */
#define PM_NOALLOC_SCRATCH_BYTES (64LL << 20)
int pm_widom_noalloc(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                     const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
                     int64_t n_per_block, int32_t precision, double* out_block_sums,
                     const pm_access* access);

#define PM_MAX_BATCH_RUNS 65536  /* runs per pm_widom_batch / pm_energy_grid_batch call */

/* One Widom run of a batch: the arguments of pm_widom. */
typedef struct pm_widom_run {
    const pm_framework* framework;
    const pm_forcefield* forcefield;
    const pm_ewald* ewald;          /* NULL when the framework or the probe is uncharged */
    const pm_probe* probe;
    double temperature;
    uint64_t seed;
    int64_t n_per_block;
    int32_t n_blocks;
    int32_t reserved;               /* 0 */
    const pm_access* access;        /* NULL: no accessibility grid */
} pm_widom_run;

/*
 * Screening: many Widom runs (frameworks x adsorbates) in one call.  The
 * reference runs `henry_coefficient` (src/henry.jl:34-181) once per
 * (crystal, molecule) pair from a user loop; this is that loop with every
 * run's insertions in the same launches (one per pipeline stage, over a work
 * list of all runs' insertions, most expensive runs first).  Each run gives
 * exactly what pm_widom gives for it; out_block_sums holds run r's n_blocks x 3
 * sums after those of runs 0..r-1.
 */
int pm_widom_batch(int32_t n_runs, const pm_widom_run* runs, int32_t precision,
                   double* out_block_sums);

/* One grid of a batch: the arguments of pm_energy_grid. */
typedef struct pm_grid_run {
    const pm_framework* framework;
    const pm_forcefield* forcefield;
    const pm_ewald* ewald;
    const pm_probe* probe;
    int32_t n_pts[3];
    int32_t reserved;               /* 0 */
    double* out;                    /* n_pts[0]*n_pts[1]*n_pts[2], i fastest */
} pm_grid_run;

/*
 * Many energy grids (src/grid.jl:291-362, one call per framework in the
 * reference) in one call: every grid's points in the same launches.  Each
 * grid equals what pm_energy_grid gives for it.
 */
int pm_energy_grid_batch(int32_t n_runs, const pm_grid_run* runs, int32_t precision);

/*
 * Point-engine tuning (no counterpart in the reference, whose CPU loops have
 * no tiling).  The chunk shapes only schedule work: every energy is a
 * function of the point and its run alone, so no setting changes a result.
 *   PM_TUNE_CELL_POINTS  points in a cube whose edge sets the cross-section
 *                        of a Widom cell (a column along the third axis, or
 *                        that cube for sparse runs; default 320)
 *   PM_TUNE_CHUNK_R_MAX  largest chunk radius (A) served by the image-list
 *                        path; wider chunks take the all-atoms path (default 6)
 * Settings are process-wide.
 */
enum { PM_TUNE_CELL_POINTS = 0, PM_TUNE_CHUNK_R_MAX = 1 };
int pm_set_tuning(int32_t key, double value);
int pm_get_tuning(int32_t key, double* value);

/*
 * Autotuner: runs the given Widom run (pm_widom's arguments) on the device for
 * a grid of (cell points, chunk radius) settings, keeps the fastest as the
 * process-wide setting (a narrower chunk-radius limit only when it is more
 * than 1% faster, and a setting is timed only when it changes the run's cell
 * layout or path) and reports it and its device time (ms).
 */
int pm_autotune_widom(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, const pm_probe* pr,
                      double temperature, uint64_t seed, int32_t n_blocks, int64_t n_per_block, int32_t precision,
                      double* out_cell_points, double* out_chunk_r_max, double* out_ms);

/*
 * Measurement utility (no counterpart in the reference): copies n_bytes
 * between two device buffers `repeats` times with a plain 16-byte-vector copy
 * kernel and reports the fastest launch in ms.  It is the known-size launch
 * that scales HBM traffic counters (FETCH_SIZE, WRITE_SIZE) of other kernels.
 */
int pm_copy_probe(int64_t n_bytes, int32_t repeats, double* out_ms);

/* Timing of the last pm_widom / pm_energy_grid / pm_vdw_energy /
 * pm_electrostatic_energy call (HIP events on the library stream): the whole
 * device pipeline of the call in ms, and the site x replicated-atom pair
 * count of the all-pairs formulation it covers. */
int pm_last_kernel_stats(double* kernel_ms, double* pair_evals);

/* For the last pm_widom / pm_energy_grid call: time of the pair kernel
 * alone (summed over batches, ms) and the (point, candidate image) pairs it
 * evaluated after spatial pruning. */
int pm_last_pair_kernel_stats(double* pair_kernel_ms, double* candidate_pairs);

/* ------------------------------------------------------------------ GCMC --
 * Grand-canonical Monte Carlo (`μVT_sim`, src/gcmc.jl:175-817) as many
 * independent Markov chains on the device: one workgroup per chain, every
 * chain with its own temperature and fugacities, all chains advanced by the
 * same launches (isotherm points and replicas run side by side).  Moves,
 * acceptance rules, proposal probabilities, the adaptive translation step and
 * the block statistics are the reference's (src/gcmc.jl:404-727,
 * src/mc_helpers.jl:17-190); energies are its guest-host and guest-guest
 * Lennard-Jones sums with the nearest image and overlap -> +Inf
 * (src/vdw_energetics.jl:25-47,85-120, src/gcmc.jl:111-124), in FP64.
 * Electrostatics are not part of the reference's GCMC (src/gcmc.jl:378-382).
 * Random numbers: move t of chain c draws 10 uniforms from Philox4x32-10 keyed
 * by the seed (the reference draws from Julia's global RNG).
 */
#define PM_MAX_GCMC_SPECIES 4
#define PM_GCMC_BLOCKS 5   /* N_BLOCKS, src/mc_helpers.jl:11 */
#define PM_GCMC_MOVES 6    /* insertion, deletion, translation, rotation, reinsertion, identity change */

typedef struct pm_gcmc pm_gcmc;

typedef struct pm_gcmc_settings {
    int64_t n_burn_cycles;       /* src/gcmc.jl:202-205 */
    int64_t n_sample_cycles;     /* >= PM_GCMC_BLOCKS */
    int32_t sample_frequency;    /* sample every this many moves (production) */
    int32_t snapshot_frequency;  /* density-grid update every this many cycles (production) */
    int32_t capacity;            /* molecules per species per chain */
    int32_t density_species;     /* adsorbate species tracked by the density grid, -1: no grid */
    uint32_t density_site_mask;  /* sites of that species counted (bit k: site k) */
    int32_t density_n_pts[3];
    int32_t density_sim_box;     /* 1: grid over the simulation box; 0: over the unit cell */
    double delta;                /* initial half translation step, A (2.0 in the reference) */
    uint64_t seed;
} pm_gcmc_settings;

typedef struct pm_gcmc_state {
    int64_t cycles_done;         /* outer cycles completed */
    int64_t markov_time;         /* moves proposed */
    int64_t num_snapshots;
    int32_t n[PM_MAX_GCMC_SPECIES];
    int32_t block;               /* current statistics block, 0-based */
    int32_t error;               /* 1: a species reached the capacity */
    double delta;
    double energy_gh, energy_gg; /* running system energy (K) */
    int64_t n_proposed[PM_GCMC_MOVES], n_accepted[PM_GCMC_MOVES];
    int64_t block_samples[PM_GCMC_BLOCKS];
    int64_t block_n[PM_GCMC_BLOCKS][PM_MAX_GCMC_SPECIES];
    int64_t block_n2[PM_GCMC_BLOCKS][PM_MAX_GCMC_SPECIES];
    double block_gh[PM_GCMC_BLOCKS], block_gg[PM_GCMC_BLOCKS];
    double block_gh2[PM_GCMC_BLOCKS], block_gg2[PM_GCMC_BLOCKS];
    double block_un[PM_GCMC_BLOCKS];
} pm_gcmc_state;

/*
 * n_species templates (probes: their LJ spheres are the molecule; charges are
 * ignored), n_chains chains with temperatures[c] (K) and fugacities[c *
 * n_species + s] (Pa).  The framework handle is the replicated simulation box
 * (`replicate(xtal, repfactors)`, src/gcmc.jl:262-264).  Chains start empty.
 */
int pm_gcmc_create(const pm_framework* fw, const pm_forcefield* ff, int32_t n_species,
                   const pm_probe* const* templates, int32_t n_chains, const double* temperatures,
                   const double* fugacities, const pm_gcmc_settings* settings, pm_gcmc** out);
int pm_gcmc_destroy(pm_gcmc* gc);
/* Starting configuration of one species of one chain (before the first
 * pm_gcmc_run): n molecules, site_xf n x n_sites x 3 (the template's LJ
 * spheres in order) and com_xf n x 3 (centres of mass), fractional in the
 * simulation box; the system energy is recomputed before the next run
 * (src/gcmc.jl:386-402). */
int pm_gcmc_set_molecules(pm_gcmc* gc, int32_t chain, int32_t species, int32_t n, const double* site_xf,
                          const double* com_xf);
/* Advance every chain by n_cycles outer cycles (each of N + 1 moves,
 * src/gcmc.jl:460-463). */
int pm_gcmc_run(pm_gcmc* gc, int64_t n_cycles);
int pm_gcmc_get_state(const pm_gcmc* gc, int32_t chain, pm_gcmc_state* out);
/* Current molecules of one species: *n of them (at most max_n copied), as
 * for pm_gcmc_set_molecules. */
int pm_gcmc_get_molecules(const pm_gcmc* gc, int32_t chain, int32_t species, int32_t max_n, int32_t* n,
                          double* site_xf, double* com_xf);
/* System energy recomputed from scratch (`total_vdw_energy`, src/gcmc.jl:735-738). */
int pm_gcmc_total_energy(pm_gcmc* gc, int32_t chain, double* out_gh, double* out_gg);
/* Density-grid counts (`update_density!`, src/grid.jl:78-106), i fastest. */
int pm_gcmc_get_density(const pm_gcmc* gc, int32_t chain, double* out);
/* No counterpart in the reference (a schedule setting, like pm_set_tuning):
 * the most molecules per species a launch keeps in LDS.  -1 (the default):
 * as many as fit beside the framework; 0: molecules stay in device memory.
 * A chain that could outgrow the LDS store within its next cycle finishes
 * the call with its molecules in device memory.  Results never depend on
 * this setting. */
int pm_gcmc_set_lds_molecules(pm_gcmc* gc, int32_t max_per_species);

#ifdef __cplusplus
}
#endif

#endif /* PMHIP_H */
