// pm_gcmc.h -- grand-canonical Monte Carlo chains of libpmhip.so (gfx950).
//
// One workgroup runs one Markov chain (src/gcmc.jl:175-817): every work-item
// draws the same random numbers and takes the same branches, so the chain's
// scalars (molecule counts, step size, energies, counters) live in registers
// of every work-item, and the work-items share out the pair sums of each
// energy evaluation.  A launch advances every chain by a number of outer
// cycles; chain state, molecules and statistics stay in device memory
// between launches.
//
// Energies follow the reference's vdw_energy (src/vdw_energetics.jl:25-47):
// fractional difference, nearest image by round-to-even, Cartesian r^2
// through the upper-triangular f_to_c, r^2 < r_c^2, overlap (r^2 < 0.1) ->
// +Inf, LJ 4 eps ratio (ratio - 1), ratio = (sigma^2/r^2)^3; guest-host over
// the replicated framework and guest-guest over every other molecule
// (src/gcmc.jl:111-124, src/vdw_energetics.jl:85-120), all FP64.
#pragma once

#include <type_traits>

#include "pm_device.h"

namespace pmk {

constexpr int GC_WG = 256;
constexpr int GC_NS = PM_MAX_GCMC_SPECIES;
constexpr int GC_NB = PM_GCMC_BLOCKS;
constexpr int GC_NM = PM_GCMC_MOVES;
enum { MV_INSERTION = 0, MV_DELETION = 1, MV_TRANSLATION = 2, MV_ROTATION = 3, MV_REINSERTION = 4, MV_IDENTITY = 5 };
constexpr double GC_KB = 1.38064852e7;  // Pa A^3 / K, src/PorousMaterials.jl:14

// Shared by every chain of a handle (device memory).
struct GcmcParams {
    double f[9];             // f_to_c of the simulation box, column-major
    double c2f[9];           // c_to_f, column-major
    double r2_cut, omega;    // LJ cutoff^2 (A^2), simulation box volume (A^3)
    double rep[3];           // replication of the unit cell (density grid over the unit cell)
    int n_atoms;             // replicated framework atoms
    const double4* atoms;    // fractional xyz + compact framework type in w
    // cell lists: the atoms within r_c + the cell's half diagonal of each cell
    // of an n_cell[0] x n_cell[1] x n_cell[2] grid over the box (null: none)
    int n_cell[3];
    const int* cell_start;   // [cells + 1]
    const int* cell_atoms;   // atom indices, cell after cell
    int n_species, n_ads_types, capacity;
    int n_sites[GC_NS];
    int site_prefix[GC_NS];  // sites of the species before each one
    int total_sites;         // sites of one molecule of every species
    int rotatable[GC_NS];
    int site_type[GC_NS][PM_MAX_SITES];
    double off[GC_NS][PM_MAX_SITES][3];  // Cartesian offsets from the centre of mass
    const double2* lj_fg;    // [framework type][adsorbate type] (sigma^2, eps)
    const double2* lj_gg;    // [adsorbate type][adsorbate type]
    int n_fw_types;          // rows of lj_fg
    double cum[GC_NM];       // cumulative proposal probabilities (last used entry 2.0)
    long long n_burn, n_sample, adjust_every, cycles_per_block;
    int sample_frequency, snapshot_frequency;
    int dens_species, dens_sim_box;
    unsigned dens_mask;
    int dens_n[3];
    unsigned long long seed;
};

// One chain (device memory).
struct GcmcChain {
    double temperature;
    double fug[GC_NS];       // Pa
    double delta, e_gh, e_gg;
    long long outer, t, n_trans_prop, n_trans_acc, num_snapshots;
    int n[GC_NS];
    int block, error;
    long long n_prop[GC_NM], n_acc[GC_NM];
    long long blk_samples[GC_NB];
    long long blk_n[GC_NB][GC_NS], blk_n2[GC_NB][GC_NS];
    double blk_gh[GC_NB], blk_gg[GC_NB], blk_gh2[GC_NB], blk_gg2[GC_NB], blk_un[GC_NB];
    double4* sites;          // [species][capacity][PM_MAX_SITES] fractional xyz
    double4* com;            // [species][capacity]
    unsigned* density;       // counts per grid point, or null
    long long outer_target;  // outer cycle the current pm_gcmc_run call ends at
};

// Where a chain's molecules live during a launch.  GcMolsMem: device memory
// in the handle's layout ([species][capacity][PM_MAX_SITES] sites,
// [species][capacity] centres).  GcMolsLds: LDS, compact per species
// ([cap][n_sites] sites of species s after those of the species before it),
// the centres ([species][cap]) after all sites.
struct GcMolsMem {
    double4* sites;
    double4* coms;
    int cap;
    __device__ double4* site_base(const GcmcParams& p, int s) const { return sites + s * cap * PM_MAX_SITES; }
    __device__ int stride(const GcmcParams& p, int s) const { return PM_MAX_SITES; }
    __device__ double4* com_base(const GcmcParams& p, int s) const { return coms + s * cap; }
    __device__ double4& site(const GcmcParams& p, int s, int m, int k) const {
        return site_base(p, s)[m * PM_MAX_SITES + k];
    }
    __device__ double4& com(const GcmcParams& p, int s, int m) const { return com_base(p, s)[m]; }
};

struct GcMolsLds {
    double4* base;
    int cap;
    __device__ double4* site_base(const GcmcParams& p, int s) const { return base + p.site_prefix[s] * cap; }
    __device__ int stride(const GcmcParams& p, int s) const { return p.n_sites[s]; }
    __device__ double4* com_base(const GcmcParams& p, int s) const { return base + (p.total_sites + s) * cap; }
    __device__ double4& site(const GcmcParams& p, int s, int m, int k) const {
        return site_base(p, s)[m * p.n_sites[s] + k];
    }
    __device__ double4& com(const GcmcParams& p, int s, int m) const { return com_base(p, s)[m]; }
};

// Copies the stored molecules of every species from one store to another.
template <typename From, typename To>
__device__ inline void gc_mols_copy(const GcmcParams& p, const From& from, const To& to, const int* n) {
    for (int s = 0; s < p.n_species; ++s) {
        const int ns = p.n_sites[s];
        for (int j = threadIdx.x; j < n[s] * ns; j += GC_WG) {
            const int m = j / ns, k = j - m * ns;
            to.site(p, s, m, k) = from.site(p, s, m, k);
        }
        for (int m = threadIdx.x; m < n[s]; m += GC_WG) to.com(p, s, m) = from.com(p, s, m);
    }
}

// The 10 uniforms of move t of a chain: species, move, molecule, position or
// step (3), rotation (3), acceptance.
__device__ inline void gcmc_uniforms(unsigned long long seed, unsigned chain, unsigned long long t, double u[10]) {
#pragma unroll
    for (uint32_t w = 0; w < 5; ++w) {
        uint32_t c[4] = {(uint32_t)t, (uint32_t)(t >> 32), chain, 0x47434D00u | w};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        u[2 * w] = u53(c[0], c[1]);
        u[2 * w + 1] = u53(c[2], c[3]);
    }
}

// The same uniforms for GC_BATCH consecutive moves at once: lane l < 60 of
// every wave holds the Philox block (t0 + l / 5, word l % 5), so one pass of
// integer work serves 12 moves; a move reads its five blocks with readlane.
constexpr int GC_BATCH = 12;
struct GcRng {
    unsigned long long t0;
    uint32_t c[4];
};

__device__ inline void gc_rng_fill(GcRng& r, unsigned long long seed, unsigned chain, unsigned long long t0) {
    const uint32_t l = (uint32_t)min((int)(threadIdx.x & 63), 5 * GC_BATCH - 1);
    const unsigned long long t = t0 + l / 5;
    r.c[0] = (uint32_t)t;
    r.c[1] = (uint32_t)(t >> 32);
    r.c[2] = chain;
    r.c[3] = 0x47434D00u | (l % 5);
    philox4x32_10(r.c, (uint32_t)seed, (uint32_t)(seed >> 32));
    r.t0 = t0;
}

__device__ inline void gc_rng_get(const GcRng& r, unsigned long long t, double u[10]) {
    const int l0 = 5 * (int)(t - r.t0);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)r.c[0], l0 + k);
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)r.c[1], l0 + k);
        const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)r.c[2], l0 + k);
        const uint32_t f = (uint32_t)__builtin_amdgcn_readlane((int)r.c[3], l0 + k);
        u[2 * k] = u53(a, b);
        u[2 * k + 1] = u53(e, f);
    }
}

// Sum over the 64 lanes of a wave by DPP row shifts and row broadcasts
// (the total lands in lane 63), returned to every lane.
template <int CTRL, int ROW_MASK>
__device__ inline double gc_dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ inline double gc_wave_total(double v) {
    v += gc_dpp<0x111, 0xf>(v);  // row_shr:1
    v += gc_dpp<0x112, 0xf>(v);  // row_shr:2
    v += gc_dpp<0x114, 0xf>(v);  // row_shr:4
    v += gc_dpp<0x118, 0xf>(v);  // row_shr:8  -> lane 15 of each row holds the row's sum
    v += gc_dpp<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += gc_dpp<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3 -> lane 63 holds the total
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// sigma^2 / r^2 through the hardware reciprocal and two Newton steps (within
// an ulp of the quotient), then 4 eps ratio (ratio - 1), ratio = x^3.
__device__ inline double lj_pair(double r2, double2 se) {
    double inv = __builtin_amdgcn_rcp(r2);
    inv = fma(fma(-r2, inv, 1.0), inv, inv);
    inv = fma(fma(-r2, inv, 1.0), inv, inv);
    const double x = se.x * inv;
    const double ratio = x * x * x;
    return 4.0 * se.y * ratio * (ratio - 1.0);
}

// Accumulates the LJ pair (x site, y other) with the nearest image.
__device__ inline void gc_pair(const GcmcParams& p, double4 x, double4 y, double2 se, double& acc, bool& ovl) {
    double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
    dx -= rint(dx);
    dy -= rint(dy);
    dz -= rint(dz);
    const double cx = p.f[0] * dx + p.f[3] * dy + p.f[6] * dz;
    const double cy = p.f[1] * dx + p.f[4] * dy + p.f[7] * dz;
    const double cz = p.f[2] * dx + p.f[5] * dy + p.f[8] * dz;
    const double r2 = cx * cx + cy * cy + cz * cz;
    if (r2 < p.r2_cut) {
        if (r2 < R2_OVERLAP) ovl = true;
        else acc += lj_pair(r2, se);
    }
}

// Guest site jj of the flattened (species, molecule, site) list of stored
// molecules, unless it lies past the end or belongs to the excluded molecule.
template <typename Mols>
__device__ inline bool gc_guest(const GcmcParams& p, const Mols& mols, const int* n, int jj, int total,
                                int ex_s, int ex_m, double4& y, int& ty) {
    if (jj >= total) return false;
    int s2 = 0;
    while (jj >= n[s2] * p.n_sites[s2]) {
        jj -= n[s2] * p.n_sites[s2];
        ++s2;
    }
    const int ns2 = p.n_sites[s2];
    const int m2 = jj / ns2, k2 = jj - m2 * ns2;
    if (s2 == ex_s && m2 == ex_m) return false;
    y = mols.site(p, s2, m2, k2);
    ty = p.site_type[s2][k2];
    return true;
}

// Guest-host and guest-guest energy of a molecule of species ts with sites
// trial[0..n_sites), against the framework and every stored molecule except
// (ex_s, ex_m).  Each work-item sums its atoms in index order, then its guest
// sites; waves reduce by DPP and the four wave sums are added in wave order.
// +Inf on overlap.  `red` alternates between two halves (`par`), so the next
// evaluation needs no barrier before writing its partial sums.
template <typename Mols>
__device__ __attribute__((always_inline)) inline void gc_energy(const GcmcParams& p, const Mols& mols,
                                                                const double4* atoms, const double2* fg,
                                                                const double2* ggt, const int* n,
                                                                const double4* trial, int ts, int ex_s, int ex_m,
                                                                bool cells, double* red, int& par, double& gh,
                                                                double& gg) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ns = p.n_sites[ts];
    const int na = p.n_ads_types;
    int total = 0;
#pragma unroll
    for (int q = 0; q < GC_NS; ++q) total += q < p.n_species ? n[q] * p.n_sites[q] : 0;
    // the first guest site is loaded ahead of the framework loop
    int jj = tid;
    double4 y = make_double4(0.0, 0.0, 0.0, 0.0);
    int ty = 0;
    bool have = gc_guest(p, mols, n, jj, total, ex_s, ex_m, y, ty);
    double agh = 0.0, agg = 0.0;
    bool ogh = false, ogg = false;
    for (int k = 0; k < ns; ++k) {
        const double4 x = trial[k];
        const double2* col = fg + p.site_type[ts][k];
        if (cells) {
            // the atoms of the site's cell list; the others are beyond r_c
            const double w0 = x.x - floor(x.x), w1 = x.y - floor(x.y), w2 = x.z - floor(x.z);
            const int c0 = min(p.n_cell[0] - 1, (int)(w0 * p.n_cell[0]));
            const int c1 = min(p.n_cell[1] - 1, (int)(w1 * p.n_cell[1]));
            const int c2 = min(p.n_cell[2] - 1, (int)(w2 * p.n_cell[2]));
            const int cell = __builtin_amdgcn_readfirstlane(c0 + p.n_cell[0] * (c1 + p.n_cell[1] * c2));
            const int e = p.cell_start[cell + 1];
            int j = p.cell_start[cell] + tid;
            for (; j + 3 * GC_WG < e; j += 4 * GC_WG) {
                const int i0 = p.cell_atoms[j], i1 = p.cell_atoms[j + GC_WG];
                const int i2 = p.cell_atoms[j + 2 * GC_WG], i3 = p.cell_atoms[j + 3 * GC_WG];
                const double4 a0 = atoms[i0], a1 = atoms[i1], a2 = atoms[i2], a3 = atoms[i3];
                gc_pair(p, x, a0, col[(int)a0.w * na], agh, ogh);
                gc_pair(p, x, a1, col[(int)a1.w * na], agh, ogh);
                gc_pair(p, x, a2, col[(int)a2.w * na], agh, ogh);
                gc_pair(p, x, a3, col[(int)a3.w * na], agh, ogh);
            }
            for (; j < e; j += GC_WG) {
                const double4 at = atoms[p.cell_atoms[j]];
                gc_pair(p, x, at, col[(int)at.w * na], agh, ogh);
            }
            continue;
        }
        int a = tid;
        for (; a + 3 * GC_WG < p.n_atoms; a += 4 * GC_WG) {
            const double4 a0 = atoms[a], a1 = atoms[a + GC_WG], a2 = atoms[a + 2 * GC_WG], a3 = atoms[a + 3 * GC_WG];
            const double2 e0 = col[(int)a0.w * na], e1 = col[(int)a1.w * na];
            const double2 e2 = col[(int)a2.w * na], e3 = col[(int)a3.w * na];
            gc_pair(p, x, a0, e0, agh, ogh);
            gc_pair(p, x, a1, e1, agh, ogh);
            gc_pair(p, x, a2, e2, agh, ogh);
            gc_pair(p, x, a3, e3, agh, ogh);
        }
        for (; a < p.n_atoms; a += GC_WG) {
            const double4 at = atoms[a];
            gc_pair(p, x, at, col[(int)at.w * na], agh, ogh);
        }
    }
    while (jj < total) {
        if (have)
            for (int k = 0; k < ns; ++k) gc_pair(p, trial[k], y, ggt[p.site_type[ts][k] * na + ty], agg, ogg);
        jj += GC_WG;
        have = gc_guest(p, mols, n, jj, total, ex_s, ex_m, y, ty);
    }
    agh = gc_wave_total(agh);
    agg = gc_wave_total(agg);
    const bool bh = __ballot(ogh) != 0ull, bg = __ballot(ogg) != 0ull;
    double* r = red + par * (4 * (GC_WG / 64));
    par ^= 1;
    if (lane == 0) {
        r[wave] = agh;
        r[GC_WG / 64 + wave] = agg;
        r[2 * (GC_WG / 64) + wave] = bh ? 1.0 : 0.0;
        r[3 * (GC_WG / 64) + wave] = bg ? 1.0 : 0.0;
    }
    __syncthreads();
    double sh = 0.0, sg = 0.0;
    bool oh = false, og = false;
#pragma unroll
    for (int w = 0; w < GC_WG / 64; ++w) {
        sh += r[w];
        sg += r[GC_WG / 64 + w];
        oh |= r[2 * (GC_WG / 64) + w] != 0.0;
        og |= r[3 * (GC_WG / 64) + w] != 0.0;
    }
    gh = oh ? INFINITY : sh;
    gg = og ? INFINITY : sg;
}

// Removes molecule m of species s, keeping the order of the others (splice!,
// src/mc_helpers.jl:58-60): chunks of work-items read, then write one
// molecule lower.
template <typename Mols>
__device__ inline void gc_remove(const GcmcParams& p, const Mols& mols, int s, int m, int n_s) {
    double4* const sb = mols.site_base(p, s);
    double4* const cb = mols.com_base(p, s);
    const int st = mols.stride(p, s);
    const int first = (m + 1) * st, last = n_s * st;
    // sites and centres move down together, one chunk of work-items each round
    for (int i0 = first, j0 = m + 1; i0 < last; i0 += GC_WG, j0 += GC_WG) {
        const int i = i0 + threadIdx.x, j = j0 + threadIdx.x;
        double4 v = make_double4(0.0, 0.0, 0.0, 0.0), c = v;
        if (i < last) v = sb[i];
        if (j < n_s) c = cb[j];
        __syncthreads();
        if (i < last) sb[i - st] = v;
        if (j < n_s) cb[j - 1] = c;
        __syncthreads();
    }
}

// Writes the trial molecule (LDS) as molecule m of species s.
template <typename Mols>
__device__ inline void gc_store(const GcmcParams& p, const Mols& mols, int s, int m, const double4* trial,
                                double4 com) {
    const int tid = threadIdx.x;
    if (tid < p.n_sites[s]) mols.site(p, s, m, tid) = trial[tid];
    if (tid == 0) mols.com(p, s, m) = com;
    __syncthreads();
}

// Rotation of site offsets: xf = com + c_to_f (R off_k) for a new molecule
// (insert_w_random_orientation!, src/mc_helpers.jl:30-49).
__device__ inline double4 gc_new_site(const GcmcParams& p, int s, int k, double4 com, const double* R, bool rot) {
#pragma clang fp contract(off)
    double o0 = p.off[s][k][0], o1 = p.off[s][k][1], o2 = p.off[s][k][2];
    if (rot) {
        const double r0 = R[0] * o0 + R[3] * o1 + R[6] * o2;
        const double r1 = R[1] * o0 + R[4] * o1 + R[7] * o2;
        const double r2 = R[2] * o0 + R[5] * o1 + R[8] * o2;
        o0 = r0; o1 = r1; o2 = r2;
    }
    const double* c = p.c2f;
    return make_double4(com.x + (c[0] * o0 + c[3] * o1 + c[6] * o2), com.y + (c[1] * o0 + c[4] * o1 + c[7] * o2),
                        com.z + (c[2] * o0 + c[5] * o1 + c[8] * o2), 0.0);
}

// Fractional-space rotation about the centre of mass:
// xf = com + (c_to_f R f_to_c)(xf - com) (random_rotation!, src/molecule.jl:263-275).
__device__ inline void gc_frac_rotation(const GcmcParams& p, const double* R, double* Rf) {
#pragma clang fp contract(off)
    double T[9];  // c_to_f R
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            T[r + 3 * c] = p.c2f[r] * R[3 * c] + p.c2f[r + 3] * R[3 * c + 1] + p.c2f[r + 6] * R[3 * c + 2];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            Rf[r + 3 * c] = T[r] * p.f[3 * c] + T[r + 3] * p.f[3 * c + 1] + T[r + 6] * p.f[3 * c + 2];
}

__device__ inline double4 gc_rotate_site(const double* Rf, double4 x, double4 com) {
#pragma clang fp contract(off)
    const double a = x.x - com.x, b = x.y - com.y, c = x.z - com.z;
    return make_double4(Rf[0] * a + Rf[3] * b + Rf[6] * c + com.x, Rf[1] * a + Rf[4] * b + Rf[7] * c + com.y,
                        Rf[2] * a + Rf[5] * b + Rf[8] * c + com.z, 0.0);
}

__device__ inline long long gc_voxel(double x, int n) {
    long long v = (long long)floor(x * (double)n);
    if (v < 0) v += n;
    else if (v >= n) v -= n;
    return v;
}

// Density-grid update of one snapshot (update_density!, src/grid.jl:78-106):
// every counted site of every molecule of the tracked species; over the unit
// cell only molecules whose centre of mass lies in it.
template <typename Mols>
__device__ inline void gc_density(const GcmcParams& p, const Mols& mols, unsigned* density, const int* n) {
    const int s = p.dens_species;
    const int ns = p.n_sites[s];
    const int n0 = p.dens_n[0], n1 = p.dens_n[1], n2 = p.dens_n[2];
    for (int j = threadIdx.x; j < n[s] * ns; j += GC_WG) {
        const int m = j / ns, k = j - m * ns;
        if (!((p.dens_mask >> k) & 1u)) continue;
        double4 x = mols.site(p, s, m, k);
        if (!p.dens_sim_box) {
            const double4 c = mols.com(p, s, m);
            const double c0 = c.x * p.rep[0], c1 = c.y * p.rep[1], c2 = c.z * p.rep[2];
            if (c0 > 1.0 || c1 > 1.0 || c2 > 1.0 || c0 < 0.0 || c1 < 0.0 || c2 < 0.0) continue;
            x = make_double4(x.x * p.rep[0], x.y * p.rep[1], x.z * p.rep[2], 0.0);
        }
        const long long v = gc_voxel(x.x, n0) + (long long)n0 * (gc_voxel(x.y, n1) + (long long)n1 * gc_voxel(x.z, n2));
        atomicAdd(&density[v], 1u);
    }
    __syncthreads();
}

// Advances every chain (one workgroup each) to its target outer cycle: the
// chain's current cycle plus n_cycles, or with `resume` the target a
// previous launch of this call recorded.  With LDS_ATOMS the replicated
// framework is staged in LDS for the launch; with LDS_MOLS the chain's
// molecules are too (up to lds_cap per species), and a chain whose next
// cycle could outgrow them stops at that cycle's start, counts itself in
// *incomplete and leaves the rest of the call to a device-memory launch.
template <bool LDS_ATOMS, bool LDS_MOLS>
__global__ __launch_bounds__(GC_WG) void gcmc_kernel(const GcmcParams* __restrict__ pp, GcmcChain* chains,
                                                     long long n_cycles, int resume, int lds_cap,
                                                     int* incomplete) {
    // dynamic LDS: LJ tables (framework x adsorbate, adsorbate x adsorbate),
    // then with LDS_ATOMS the replicated framework, then with LDS_MOLS the
    // molecules
    extern __shared__ double4 s_dyn[];
    __shared__ double4 s_trial[PM_MAX_SITES];
    __shared__ double4 s_old[PM_MAX_SITES];
    __shared__ double s_red[2 * 4 * (GC_WG / 64)];
    const GcmcParams& p = *pp;
    GcmcChain& ch = chains[blockIdx.x];
    const int tid = threadIdx.x;
    if (ch.error) return;
    const int n_fg = p.n_fw_types * p.n_ads_types, n_gg = p.n_ads_types * p.n_ads_types;
    double2* s_fg = reinterpret_cast<double2*>(s_dyn);
    double2* s_gg = s_fg + n_fg;
    for (int i = tid; i < n_fg; i += GC_WG) s_fg[i] = p.lj_fg[i];
    for (int i = tid; i < n_gg; i += GC_WG) s_gg[i] = p.lj_gg[i];
    const double4* atoms = p.atoms;
    double4* lds_next = s_dyn + (n_fg + n_gg + 1) / 2;
    if constexpr (LDS_ATOMS) {
        double4* s_atoms = lds_next;
        for (int a = tid; a < p.n_atoms; a += GC_WG) s_atoms[a] = p.atoms[a];
        atoms = s_atoms;
        lds_next += p.n_atoms;
    }
    const GcMolsMem gmols{ch.sites, ch.com, p.capacity};
    using Mols = typename std::conditional<LDS_MOLS, GcMolsLds, GcMolsMem>::type;
    Mols mols;
    if constexpr (LDS_MOLS) mols = GcMolsLds{lds_next, lds_cap};
    else mols = gmols;
    double fug[GC_NS];
#pragma unroll
    for (int q = 0; q < GC_NS; ++q) fug[q] = ch.fug[q];
    int par = 0;
    const bool cells_on = p.cell_start != nullptr;
    const unsigned chain = blockIdx.x;
    const double T = ch.temperature;
    int n[GC_NS];
#pragma unroll
    for (int s = 0; s < GC_NS; ++s) n[s] = ch.n[s];
    if constexpr (LDS_MOLS) {
        // a chain already over the LDS capacity leaves the call to device memory
        bool fits = true;
        for (int s = 0; s < GC_NS; ++s) fits &= n[s] <= lds_cap;
        if (fits) gc_mols_copy(p, gmols, mols, n);
    }
    double delta = ch.delta, e_gh = ch.e_gh, e_gg = ch.e_gg;
    long long t = ch.t, outer = ch.outer, ntp = ch.n_trans_prop, nta = ch.n_trans_acc, nsnap = ch.num_snapshots;
    const long long target = resume ? ch.outer_target : outer + n_cycles;
    const long long outer_start = outer;
    bool stopped = false;
    long long n_prop[GC_NM], n_acc[GC_NM];
#pragma unroll
    for (int k = 0; k < GC_NM; ++k) { n_prop[k] = ch.n_prop[k]; n_acc[k] = ch.n_acc[k]; }
    int block = ch.block, error = 0;
    const int S = p.n_species;
    // countdowns to the next multiple of each period (no 64-bit division per move)
    long long to_sample = p.sample_frequency - t % p.sample_frequency;
    long long to_snap = p.snapshot_frequency - outer % p.snapshot_frequency;
    long long to_adjust = p.adjust_every > 0 ? p.adjust_every - outer % p.adjust_every : 0;
    long long to_block = 0;  // cycles to the next block boundary once sampling
    if (outer >= p.n_burn) to_block = p.cycles_per_block - (outer - p.n_burn) % p.cycles_per_block;
    GcRng rng;
    gc_rng_fill(rng, p.seed, chain, (unsigned long long)t + 1);
    // sums of the current statistics block, added to the chain's block at a
    // block change and at the end of the launch
    long long acc_samples = 0, acc_n[GC_NS], acc_n2[GC_NS];
    double acc_gh = 0.0, acc_gg = 0.0, acc_gh2 = 0.0, acc_gg2 = 0.0, acc_un = 0.0;
#pragma unroll
    for (int q = 0; q < GC_NS; ++q) { acc_n[q] = 0; acc_n2[q] = 0; }
    auto flush = [&]() {
        if (tid == 0 && acc_samples > 0) {
            ch.blk_samples[block] += acc_samples;
            for (int q = 0; q < GC_NS; ++q) {
                ch.blk_n[block][q] += acc_n[q];
                ch.blk_n2[block][q] += acc_n2[q];
            }
            ch.blk_gh[block] += acc_gh;
            ch.blk_gg[block] += acc_gg;
            ch.blk_gh2[block] += acc_gh2;
            ch.blk_gg2[block] += acc_gg2;
            ch.blk_un[block] += acc_un;
        }
        acc_samples = 0;
        acc_gh = acc_gg = acc_gh2 = acc_gg2 = acc_un = 0.0;
        for (int q = 0; q < GC_NS; ++q) { acc_n[q] = 0; acc_n2[q] = 0; }
    };
    __syncthreads();

    while (outer < target && !error) {
        int n_tot = 0;
        for (int s = 0; s < S; ++s) n_tot += n[s];
        if constexpr (LDS_MOLS) {
            // a cycle of n_tot + 1 moves adds at most n_tot + 1 molecules to a species
            bool fits = true;
            for (int s = 0; s < S; ++s) fits &= n[s] + n_tot + 1 <= lds_cap;
            if (!fits) {
                stopped = true;
                break;
            }
        }
        ++outer;
        const long long inner = (long long)n_tot + 1;
        for (long long it = 0; it < inner; ++it) {
            ++t;
            if ((unsigned long long)t >= rng.t0 + GC_BATCH) gc_rng_fill(rng, p.seed, chain, (unsigned long long)t);
            double u[10];
            gc_rng_get(rng, (unsigned long long)t, u);
            const int s = S > 1 ? min(S - 1, (int)(u[0] * (double)S)) : 0;
            const int n_i = n[s];
            int mv = 0;
            while (mv < GC_NM - 1 && !(u[1] < p.cum[mv])) ++mv;
            n_prop[mv] += 1;
            const int ns = p.n_sites[s];
            const bool rot = p.rotatable[s] != 0;
            const int m = n_i > 0 ? min(n_i - 1, (int)(u[2] * (double)n_i)) : 0;
            if (mv == MV_INSERTION) {
                if (n_i == p.capacity) {
                    error = 1;
                    break;
                }
                double R[9];
                if (rot) arvo_rotation(&u[6], R);
                const double4 com = make_double4(u[3], u[4], u[5], 0.0);
                if (tid < ns) s_trial[tid] = gc_new_site(p, s, tid, com, R, rot);
                __syncthreads();
                double gh, gg;
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_trial, s, -1, -1, cells_on, s_red, par, gh, gg);
                const double dE = (gg + 0.0) + gh + 0.0;
                const double acc = fug[s] * p.omega / ((double)(n_i + 1) * GC_KB * T) * exp(-dE / T);
                if (u[9] < acc) {
                    gc_store(p, mols, s, n_i, s_trial, com);
                    n[s] = n_i + 1;
                    n_acc[mv] += 1;
                    e_gh += gh;
                    e_gg += gg;
                }
            } else if (mv == MV_DELETION && n_i != 0) {
                if (tid < ns) s_trial[tid] = mols.site(p, s, m, tid);
                __syncthreads();
                double gh, gg;
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_trial, s, s, m, cells_on, s_red, par, gh, gg);
                const double dE = (gg + 0.0) + gh + 0.0;
                const double acc = (double)n_i * GC_KB * T / (fug[s] * p.omega) * exp(dE / T);
                if (u[9] < acc) {
                    gc_remove(p, mols, s, m, n_i);
                    n[s] = n_i - 1;
                    n_acc[mv] += 1;
                    e_gh -= gh;
                    e_gg -= gg;
                }
            } else if ((mv == MV_TRANSLATION || (mv == MV_ROTATION && rot) || mv == MV_REINSERTION) && n_i != 0) {
                if (mv == MV_TRANSLATION) ntp += 1;
                const double4 com0 = mols.com(p, s, m);
                if (tid < ns) s_old[tid] = mols.site(p, s, m, tid);
                __syncthreads();
                double gh0, gg0, gh1, gg1;
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_old, s, s, m, cells_on, s_red, par, gh0, gg0);
                double4 com1 = com0;
                {
#pragma clang fp contract(off)
                    if (mv == MV_TRANSLATION) {
                        // Cartesian step in a cube of side delta, then periodic COM
                        // (random_translation!, src/mc_helpers.jl:142-153)
                        const double d0 = delta * (u[3] - 0.5), d1 = delta * (u[4] - 0.5), d2 = delta * (u[5] - 0.5);
                        const double* c = p.c2f;
                        const double f0 = c[0] * d0 + c[3] * d1 + c[6] * d2;
                        const double f1 = c[1] * d0 + c[4] * d1 + c[7] * d2;
                        const double f2 = c[2] * d0 + c[5] * d1 + c[8] * d2;
                        double4 x = tid < ns ? s_old[tid] : make_double4(0.0, 0.0, 0.0, 0.0);
                        x.x += f0; x.y += f1; x.z += f2;
                        com1.x += f0; com1.y += f1; com1.z += f2;
                        if (com1.x > 1.0 || com1.y > 1.0 || com1.z > 1.0 || com1.x < 0.0 || com1.y < 0.0 ||
                            com1.z < 0.0) {
                            // apply_periodic_boundary_condition!, src/mc_helpers.jl:68-89
                            double4 nc = com1;
                            nc.x = com1.x >= 1.0 ? com1.x - 1.0 : (com1.x < 0.0 ? com1.x + 1.0 : com1.x);
                            nc.y = com1.y >= 1.0 ? com1.y - 1.0 : (com1.y < 0.0 ? com1.y + 1.0 : com1.y);
                            nc.z = com1.z >= 1.0 ? com1.z - 1.0 : (com1.z < 0.0 ? com1.z + 1.0 : com1.z);
                            const double e0 = nc.x - com1.x, e1 = nc.y - com1.y, e2 = nc.z - com1.z;
                            x.x += e0; x.y += e1; x.z += e2;
                            com1.x += e0; com1.y += e1; com1.z += e2;
                        }
                        if (tid < ns) s_trial[tid] = x;
                    } else {
                        double4 x = tid < ns ? s_old[tid] : make_double4(0.0, 0.0, 0.0, 0.0);
                        if (mv == MV_REINSERTION) {
                            // translate_to! a uniform point (random_reinsertion!, src/mc_helpers.jl:168-182)
                            const double e0 = u[3] - com0.x, e1 = u[4] - com0.y, e2 = u[5] - com0.z;
                            x.x += e0; x.y += e1; x.z += e2;
                            com1.x += e0; com1.y += e1; com1.z += e2;
                        }
                        if (rot) {
                            double R[9], Rf[9];
                            arvo_rotation(&u[6], R);
                            gc_frac_rotation(p, R, Rf);
                            x = gc_rotate_site(Rf, x, com1);
                        }
                        if (tid < ns) s_trial[tid] = x;
                    }
                }
                __syncthreads();
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_trial, s, s, m, cells_on, s_red, par, gh1, gg1);
                const double e_new = (gg1 + 0.0) + gh1 + 0.0, e_old = (gg0 + 0.0) + gh0 + 0.0;
                if (u[9] < exp(-(e_new - e_old) / T)) {
                    gc_store(p, mols, s, m, s_trial, com1);
                    n_acc[mv] += 1;
                    if (mv == MV_TRANSLATION) nta += 1;
                    e_gh += gh1 - gh0;
                    e_gg += gg1 - gg0;
                }
            } else if (mv == MV_IDENTITY && n_i != 0) {
                // src/gcmc.jl:566-613
                const double4 com0 = mols.com(p, s, m);
                if (tid < ns) s_old[tid] = mols.site(p, s, m, tid);
                __syncthreads();
                double gh0, gg0, gh1, gg1;
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_old, s, s, m, cells_on, s_red, par, gh0, gg0);
                gc_remove(p, mols, s, m, n_i);
                n[s] = n_i - 1;
                int pick = min(S - 2, (int)(u[3] * (double)(S - 1)));
                const int s2 = pick < s ? pick : pick + 1;
                const int n_j = n[s2];
                if (n_j == p.capacity) {
                    error = 1;
                    break;
                }
                double R[9];
                const bool rot2 = p.rotatable[s2] != 0;
                if (rot2) arvo_rotation(&u[6], R);
                if (tid < p.n_sites[s2]) s_trial[tid] = gc_new_site(p, s2, tid, com0, R, rot2);
                __syncthreads();
                gc_energy(p, mols, atoms, s_fg, s_gg, n, s_trial, s2, -1, -1, cells_on, s_red, par, gh1, gg1);
                const double e_new = (gg1 + 0.0) + gh1 + 0.0, e_old = (gg0 + 0.0) + gh0 + 0.0;
                const double acc = ((double)n_i * fug[s2] / ((double)(n_j + 1) * fug[s])) * exp(-(e_new - e_old) / T);
                if (u[9] < acc) {
                    gc_store(p, mols, s2, n_j, s_trial, com0);
                    n[s2] = n_j + 1;
                    n_acc[mv] += 1;
                    e_gh += gh1 - gh0;
                    e_gg += gg1 - gg0;
                } else {
                    // the original molecule goes back at the end of its species' list
                    gc_store(p, mols, s, n[s], s_old, com0);
                    n[s] += 1;
                }
            }
            const bool sample_now = --to_sample == 0;
            if (sample_now) to_sample = p.sample_frequency;
            if (outer > p.n_burn && sample_now) {
                long long ntot = 0;
#pragma unroll
                for (int q = 0; q < GC_NS; ++q) {
                    ntot += n[q];
                    acc_n[q] += n[q];
                    acc_n2[q] += (long long)n[q] * n[q];
                }
                acc_samples += 1;
                acc_gh += e_gh;
                acc_gg += e_gg;
                acc_gh2 += e_gh * e_gh;
                acc_gg2 += e_gg * e_gg;
                acc_un += ((e_gg + 0.0) + e_gh + 0.0) * (double)ntot;
            }
        }
        if (error) break;
        // adaptive step during burn-in (src/gcmc.jl:634-637, src/mc_helpers.jl:107-129)
        bool adjust_now = false;
        if (p.adjust_every > 0 && --to_adjust == 0) {
            adjust_now = true;
            to_adjust = p.adjust_every;
        }
        if (outer <= p.n_burn && adjust_now) {
            const double rate = (double)nta / (double)ntp;
            if (rate > 0.5) delta = 1.2 * delta;
            else if (rate < 0.4) delta = delta / 1.2;
            ntp = 0;
            nta = 0;
        }
        // statistics blocks (src/gcmc.jl:639-653)
        bool block_now = false;
        if (outer > p.n_burn) {
            if (to_block == 0) to_block = p.cycles_per_block;  // first sampling cycle
            if (--to_block == 0) block_now = true;
        }
        if (outer > p.n_burn && block != GC_NB - 1 && block_now) {
            flush();
            ++block;
        }
        // snapshots (src/gcmc.jl:663-678)
        const bool snap_now = --to_snap == 0;
        if (snap_now) to_snap = p.snapshot_frequency;
        if (outer > p.n_burn && snap_now) {
            if (p.dens_species >= 0) gc_density(p, mols, ch.density, n);
            ++nsnap;
        }
    }
    flush();
    __syncthreads();
    if constexpr (LDS_MOLS) {
        bool fits = true;
        for (int s = 0; s < GC_NS; ++s) fits &= n[s] <= lds_cap;
        if (fits && outer != outer_start) gc_mols_copy(p, mols, gmols, n);
        if (stopped && tid == 0) atomicAdd(incomplete, 1);
    }
    if (tid == 0) {
        ch.outer_target = target;
        for (int s = 0; s < GC_NS; ++s) ch.n[s] = n[s];
        ch.delta = delta;
        ch.e_gh = e_gh;
        ch.e_gg = e_gg;
        ch.t = t;
        ch.outer = outer;
        ch.n_trans_prop = ntp;
        ch.n_trans_acc = nta;
        ch.num_snapshots = nsnap;
        for (int k = 0; k < GC_NM; ++k) { ch.n_prop[k] = n_prop[k]; ch.n_acc[k] = n_acc[k]; }
        ch.block = block;
        ch.error = error;
    }
}

// System energy from scratch (total_vdw_energy, src/vdw_energetics.jl:146-193):
// guest-host summed over molecules, guest-guest summed over molecules and halved.
__global__ __launch_bounds__(GC_WG) void gcmc_total_energy_kernel(const GcmcParams* __restrict__ pp,
                                                                  GcmcChain* chains, double* out) {
    __shared__ double4 s_trial[PM_MAX_SITES];
    __shared__ double s_red[2 * 4 * (GC_WG / 64)];
    const GcmcParams& p = *pp;
    GcmcChain& ch = chains[blockIdx.x];
    const GcMolsMem mols{ch.sites, ch.com, p.capacity};
    int n[GC_NS];
#pragma unroll
    for (int s = 0; s < GC_NS; ++s) n[s] = ch.n[s];
    int par = 0;
    double tgh = 0.0, tgg = 0.0;
    for (int s = 0; s < p.n_species; ++s) {
        for (int m = 0; m < n[s]; ++m) {
            if (threadIdx.x < p.n_sites[s]) s_trial[threadIdx.x] = mols.site(p, s, m, threadIdx.x);
            __syncthreads();
            double gh, gg;
            gc_energy(p, mols, p.atoms, p.lj_fg, p.lj_gg, n, s_trial, s, s, m, false, s_red, par, gh, gg);
            tgh += gh;
            tgg += gg;
        }
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = tgh;
        out[2 * blockIdx.x + 1] = tgg / 2.0;
    }
}

}  // namespace pmk
