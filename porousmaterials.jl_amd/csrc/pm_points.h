// pm_points.h -- the point engine of libpmhip.so (gfx950, wave64).
//
// Every probe site is an independent point.  A call runs a work list of runs
// (Widom runs: framework, probe, insertion stream and block layout; or energy
// grids), cut into segments that fill launch groups of at most a fixed number
// of point slots.  Each stage is one launch over every segment of a group:
//
//   widom_gen_kernel        poses from the counter stream, sites wrapped into
//                           [0,1)^3, binned into the run's cells and given
//                           slots within super-bins of cells
//   scan_excl_kernel,
//   stage_points_kernel,
//   cell_sort_kernel,
//   slice_items_kernel      counting sort by cell; work items (cell, slice)
//   point_pair_kernel       LJ + real-space Ewald of a chunk of points (one
//                           slice of a cell, or one brick of a grid) against
//                           the framework images near the chunk
//   point_fourier_kernel    FP64 Fourier part per charged point
//   widom_combine_kernel    per insertion energy, Boltzmann weight, partials
//   grid_combine_kernel     per grid point energy
//   block_sums_kernel       fixed-order per-block sums of every run
//
// Per-run parameters live in device memory (RunArgs); a workgroup finds its
// segment (and with it its run) by a binary search that is uniform across the
// workgroup.  A point's energy is a function of the point and its run alone:
// pair vectors are differences of absolute Cartesian coordinates, and the
// FP32 partial sums follow the image list, which is fixed by the framework and
// the cutoff radius.  The cell layout, the segment and launch group a point
// lands in and the tuning of chunk sizes therefore never change a result.
#pragma once

#include "pm_device.h"

namespace pmk {

// Images per spatial tile of an image list (one wave loads a tile).
constexpr int TILE_IMAGES = 64;

constexpr int ERFC_SEG = 512;      // cubic segments of erfc(alpha r) on [0, r_sr] (FP32 path)
// The largest float below ERFC_SEG: the scaled distance is clamped to it, so
// the last segment keeps its own t in [0, 1) up to r_sr, a distance beyond r_sr
// (a masked term) reads a finite value, and r = 0 (0 * inf = NaN) reads
// segment ERFC_SEG - 1 rather than a NaN.
constexpr float ERFC_XMAX = (float)ERFC_SEG - (float)ERFC_SEG / 16777216.0f;
constexpr int ERFC64_SEG = 512;    // quintic segments of erfc(alpha r) on [0, r_sr] (FP64 path)
constexpr int SCAN_WG = 1024;      // work-items of the scan kernels
constexpr int GT = PM_MAX_GRAD_TYPES;  // framework types of a gradient run (fixed stride)
constexpr int GQ = 4;                  // per-type sums per point: d/d eps, d/d sigma, erfc/r, Fourier
constexpr int GK = 3 + 3 * GT;         // per-workgroup partials of a gradient run
static_assert(7 * GT <= 64, "a tile's cumulative type counts (at most 64, 7 bits each) fill one 64-bit word");

// One Widom run or grid of a call (device memory).
struct RunArgs {
    // replicated box
    double f11, f12, f13, f22, f23, f33;  // f_to_c (upper triangular)
    double c2f[9];                        // c_to_f, column-major
    double repd[3];
    // pair terms
    double r2_lj, r2_sr, alpha, r_cmax, erfc_inv_h;
    double chunk_r;          // bound on |point - chunk centre| (Cartesian)
    double inv4pe;
    int esm;                 // 0 no electrostatics, 1 LJ and Ewald share r_c, 2 separate cutoffs
    int fast;                // the image list covers r_cmax + chunk_r around every chunk
    int nsite, n_types, n_img, n_atoms, n_tiles;
    unsigned site_lj_mask, site_q_mask;
    const double4* img;      // images near the box: Cartesian xyz (A) + charge, in spatial tiles
    const int* img_type;
    const double4* tile_sph; // per tile: bounding-sphere centre (Cartesian) and radius
    const int2* tile_rng;    // per tile: first image, image count (<= 64)
    const double4* atoms;    // replicated atoms (fallback path)
    const int* species;
    const void* ljtab;       // [n_types][nsite] LJP<Real>
    const float4* erfc_tab;  // ERFC_SEG cubics of erfc(alpha r) in t = r/h - i
    const double2* erfc64_tab;  // ERFC64_SEG quintics (c0, c1), (c2, c3), (c4, c5) in t = r/h - i
    double erfc64_inv_h;
    // Fourier box: ma in [0, Ma], mb in [-Mb, Mb], mc in [-Mc, Mc]
    int Ma, Mb, Mc;
    const double2* A;        // wt * N_R * S(m), zeros off the kept set
    const int2* krows;       // per (ma, mb) row: kept mc range (lo > hi: empty row)
    // Widom
    unsigned long long seed;
    long long n_per_block;
    int rotate, n_blocks, chunks_per_block, block_base;
    int ncell[3];
    int zbins;               // slabs per cell along the third axis (points are ordered by slab)
    double off[PM_MAX_SITES][3];  // site offsets from the COM (Cartesian)
    double site_q[PM_MAX_SITES];
    double temperature;
    double* partials;        // per insertion workgroup of the run: (S, W_vdw, W_es)
    double* out_e;           // nullable: (E_vdw, E_es) per insertion
    const unsigned char* acc;  // nullable: accessibility grid over the unit cell, i fastest
    int acc_n[3];
    // grid
    int grid_n[3], brick[3], nbrick[3];
    double grid_q;
    double* grid_d;          // FP64 output, or
    float* grid_f;           // FP32 output
    // parameter gradients (adjoint Widom runs)
    int n_gt;                // framework types (<= GT)
    int pstride;             // doubles per workgroup partial: 3, or GK for a gradient run
    const void* gtab;        // [n_types][nsite] Real: 4 eps_ts / sigma_ts (d LJ / d sigma_ts factor)
    const double4* jac;      // [n_types][nsite]: d eps_ts/d eps_t, d eps_ts/d sigma_t, d sigma_ts/d eps_t,
                             // d sigma_ts/d sigma_t of the mixing rule
    const double2* A_t;      // per-type Fourier boxes: GT complex coefficients per box entry
    double* gsums;           // per block: 3 GT sums of w dE/d(eps_t, sigma_t, q_t)
};

// A contiguous piece of one run inside a launch group.
struct SegArgs {
    int run;
    int w0;               // Widom: first insertion workgroup of the run in this segment
    int wg_base;          // first gen/combine workgroup (Widom) or brick (grid) of the segment
    unsigned cell_base;   // Widom: first cell of the segment
    int gen_base;         // Widom: first generator workgroup of the segment
    int pad;
    long long slot_base;  // first point slot of the segment (a multiple of WG)
    long long n_slots;    // point slots of the segment (grid: the grid's points)
};

struct GroupArgs {
    const RunArgs* runs;
    const SegArgs* segs;
    int n_seg, mode;
    double4* pos;            // Widom: per slot, fractional xyz in [0,1)^3, w = 1 (valid) or 0
    uint2* cell;             // Widom: per slot, (slab << 24 | cell in the group, or 0xffffffff;
                             // copy << 28 | slot within the super-bin copy)
    unsigned* sb_count;      // Widom: points per counter copy of each super-bin of 2^sb_shift cells
    int sb_shift;
    const unsigned* point_off;
    const uint4* items;      // (cell, slice, segment, 0)
    const unsigned* n_items;
    const unsigned* order;   // slots sorted by cell
    double* out_lj;          // per slot
    double* out_sr;
    double* out_lr;
    unsigned long long* cand_pairs;  // evaluated candidate pairs (reporting)
    double* out_g;           // gradient runs: GQ x GT per-type sums per slot, sum-major:
                             // sum (q GT + t) of slot p at out_g[(q GT + t) g_stride + p]
    long long g_stride;
};

// A pointer into read-only device memory (the constant address space: the
// kernel never writes what it points to) that the compiler may treat as
// uniform across the wave (it is: every lane holds the same value), so that
// loads through uniform indices become scalar loads.
template <typename T>
using gconst_ptr = const T __attribute__((address_space(4)))*;

template <typename T>
__device__ inline gconst_ptr<T> uniform_ptr(const T* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<gconst_ptr<T>>(((unsigned long long)hi << 32) | lo);
}

enum { SEG_BY_WG = 0, SEG_BY_CELL = 1, SEG_BY_SLOT = 2, SEG_BY_GEN = 3 };

// Last segment whose first workgroup / cell / slot is <= v.
template <int KEY>
__device__ inline int find_segment(const SegArgs* s, int n, long long v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long k = KEY == SEG_BY_WG     ? (long long)s[mid].wg_base
                            : KEY == SEG_BY_CELL ? (long long)s[mid].cell_base
                            : KEY == SEG_BY_GEN  ? (long long)s[mid].gen_base
                                                 : s[mid].slot_base;
        if (k <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline double grid_coord(double idx, int n, double rep) {
    return (n > 1 ? idx / (double)(n - 1) : 0.0) / rep;
}

// Accessibility of an insertion's centre of mass u (replicated fractional):
// unit-cell coordinates x = mod(u R, 1), the nearest grid point
// round(x (n - 1)) (ties to even) and its 26 neighbours, index -1 -> n - 1 and
// n -> 0; open if any of the 27 points is accessible (src/grid.jl:778-836).
// Contraction is off so that u R and x (n - 1) round as two plain products.
__device__ inline bool access_open(const RunArgs& a, const double* u) {
#pragma clang fp contract(off)
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double x = u[d] * a.repd[d];
        x = x - floor(x);
        c[d] = (int)rint(x * (double)(a.acc_n[d] - 1));
    }
    const int n0 = a.acc_n[0], n1 = a.acc_n[1], n2 = a.acc_n[2];
    bool open = false;
    for (int dk = -1; dk <= 1; ++dk) {
        int k = c[2] + dk;
        k = k < 0 ? n2 - 1 : (k >= n2 ? 0 : k);
        for (int dj = -1; dj <= 1; ++dj) {
            int j = c[1] + dj;
            j = j < 0 ? n1 - 1 : (j >= n1 ? 0 : j);
            for (int di = -1; di <= 1; ++di) {
                int i = c[0] + di;
                i = i < 0 ? n0 - 1 : (i >= n0 ? 0 : i);
                open |= a.acc[(long long)i + (long long)n0 * ((long long)j + (long long)n1 * k)] != 0;
            }
        }
    }
    return open;
}

// ------------------------------------------------------------ Widom gen ----
// One work-item per insertion slot: pose from the stream, sites folded into
// the first unit cell [0, 1 / rep)^3 of the replicated box (shifts by unit-cell
// vectors change neither a site's set of pair distances -- the framework has
// the unit cell's period -- nor the Fourier phases m.(R o x)), binned into the
// run's cells over that unit cell: a cell of a given point count is rep_a rep_b
// rep_c times smaller than over the whole box, and so is the slack of a
// chunk's candidate sphere beyond the cutoff.  An insertion in a
// blocked pocket of the run's accessibility grid emits no point (its slots
// look like padding); the combine kernels give it E_vdw = E_es = +Inf
// (src/henry.jl:225-232).
//
// A generator workgroup covers GEN_CHUNKS consecutive insertion workgroups of
// one segment (WG insertions each).  Binning is the first of two counting
// passes.  Cells come in super-bins of 2^sb_shift consecutive cells; the
// workgroup counts its points per super-bin in LDS and takes each super-bin's
// slots with one global atomic on one of SB_COPIES counters of the super-bin
// (workgroup index mod SB_COPIES, so that no counter serialises every
// workgroup's atomics), and a point leaves with (slab << 24 | cell,
// copy << 28 | slot within its copy of the super-bin).  One device-wide atomic
// per point would cost more than the rest of the kernel.
constexpr int GEN_CHUNKS = 4;       // insertion workgroups per generator workgroup
constexpr int SB_SPAN_MAX = 4160;   // super-bins one segment's cells may span
constexpr int SB_KEYS_MAX = 4096;   // (cell, slab) keys per super-bin: 2^sb_shift x the group's slab stride
constexpr int SB_COPIES = 16;       // counters per super-bin
constexpr unsigned SB_SLOT_MASK = 0x0fffffffu;
// Cells are columns along the third axis (cubes, with one slab, for sparse
// runs), and within a cell the points are ordered by `zbins` slabs along that
// axis, so that a slice of WG consecutive points is a compact piece of the
// column and the workgroup's candidate sphere is small; the slab rides in the
// top bits of the cell index (cells of a group < 2^24, slabs <= ZBINS_MAX).
constexpr int ZBINS_MAX = 255;
constexpr unsigned CELL_MASK = 0x00ffffffu;
__global__ __launch_bounds__(WG * GEN_CHUNKS) void widom_gen_kernel(GroupArgs g) {
    constexpr int GEN_WG = WG * GEN_CHUNKS;
    __shared__ unsigned s_sb[SB_SPAN_MAX];
    const SegArgs sg = g.segs[find_segment<SEG_BY_GEN>(g.segs, g.n_seg, blockIdx.x)];
    const RunArgs& a = g.runs[sg.run];
    const int ns = a.nsite;
    const int t = threadIdx.x % WG;
    // insertion workgroup within the segment; past the segment's last one the
    // work-item writes nothing (those slots are another segment's)
    const int lw = ((int)blockIdx.x - sg.gen_base) * GEN_CHUNKS + (int)threadIdx.x / WG;
    const bool in_seg = (long long)lw * WG * ns < sg.n_slots;
    const int w = sg.w0 + lw;  // insertion workgroup within the run
    const int b = w / a.chunks_per_block, c = w - b * a.chunks_per_block;
    const long long i = (long long)c * WG + t;
    const bool valid = in_seg && i < a.n_per_block;
    const long long p0 = sg.slot_base + ((long long)lw * WG + t) * ns;
    const int n0 = a.ncell[0], n1 = a.ncell[1], n2 = a.ncell[2];
    // the super-bins of this segment's cells
    const int S = g.sb_shift;
    const unsigned sb_lo = sg.cell_base >> S;
    const int span = (int)(((sg.cell_base + (unsigned)(n0 * n1 * n2) - 1u) >> S) - sb_lo) + 1;
    for (int k = threadIdx.x; k < span; k += GEN_WG) s_sb[k] = 0u;
    double u[6];
    bool emit = valid;
    if (valid) {
        const unsigned long long gi = (unsigned long long)b * a.n_per_block + i;
        insertion_uniforms(a.seed, gi, u);
        if (a.acc != nullptr) emit = access_open(a, u);
    }
    __syncthreads();  // the super-bin counts are clear
    unsigned cid[PM_MAX_SITES], slot[PM_MAX_SITES], zb[PM_MAX_SITES];
    if (!in_seg) {
    } else if (!emit) {
        for (int s = 0; s < ns; ++s) {
            g.pos[p0 + s] = make_double4(0.0, 0.0, 0.0, 0.0);
            g.cell[p0 + s] = make_uint2(0xffffffffu, 0u);
        }
    } else {
        double R[9];
        const bool rotate = a.rotate != 0;
        if (rotate) arvo_rotation(&u[3], R);
        const gconst_ptr<double> offp = uniform_ptr(&a.off[0][0]);
        const gconst_ptr<double> c2fp = uniform_ptr(&a.c2f[0]);
        const gconst_ptr<double> repp = uniform_ptr(&a.repd[0]);
        double c2f[9], repd[3];
#pragma unroll
        for (int k = 0; k < 9; ++k) c2f[k] = c2fp[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) repd[k] = repp[k];
#pragma unroll
        for (int s = 0; s < PM_MAX_SITES; ++s) {
            if (s >= ns) break;
            double y0 = offp[3 * s], y1 = offp[3 * s + 1], y2 = offp[3 * s + 2];
            if (rotate) {
                const double t0 = R[0] * y0 + R[3] * y1 + R[6] * y2;
                const double t1 = R[1] * y0 + R[4] * y1 + R[7] * y2;
                const double t2 = R[2] * y0 + R[5] * y1 + R[8] * y2;
                y0 = t0; y1 = t1; y2 = t2;
            }
            // unit-cell fractional coordinates in [0, 1), then back to the box's
            double x = (u[0] + (c2f[0] * y0 + c2f[3] * y1 + c2f[6] * y2)) * repd[0];
            double y = (u[1] + (c2f[1] * y0 + c2f[4] * y1 + c2f[7] * y2)) * repd[1];
            double z = (u[2] + (c2f[2] * y0 + c2f[5] * y1 + c2f[8] * y2)) * repd[2];
            x -= floor(x);
            y -= floor(y);
            z -= floor(z);
            const int ia = min((int)(x * n0), n0 - 1);
            const int ib = min((int)(y * n1), n1 - 1);
            const int ic = min((int)(z * n2), n2 - 1);
            const int iz = min((int)((z * n2 - ic) * a.zbins), a.zbins - 1);
            cid[s] = sg.cell_base + (unsigned)(ia + n0 * (ib + n1 * ic));
            zb[s] = (unsigned)iz << 24;
            g.pos[p0 + s] = make_double4(x / repd[0], y / repd[1], z / repd[2], 1.0);
            // the count before this point is its rank among the workgroup's
            // points of the super-bin (the order within a cell changes no
            // point's energy)
            slot[s] = atomicAdd(&s_sb[(cid[s] >> S) - sb_lo], 1u);
        }
    }
    __syncthreads();  // the workgroup's counts are complete
    const unsigned copy = blockIdx.x % SB_COPIES;
    for (int k = threadIdx.x; k < span; k += GEN_WG) {
        const unsigned n = s_sb[k];
        if (n != 0u) s_sb[k] = atomicAdd(&g.sb_count[(sb_lo + k) * SB_COPIES + copy], n);
    }
    __syncthreads();  // each super-bin's first slot for this workgroup
    if (emit) {
#pragma unroll
        for (int s = 0; s < PM_MAX_SITES; ++s) {
            if (s >= ns) break;
            g.cell[p0 + s] = make_uint2(zb[s] | cid[s], (copy << 28) | (s_sb[(cid[s] >> S) - sb_lo] + slot[s]));
        }
    }
}

__global__ void fill_u32_kernel(unsigned* p, long long n, unsigned v, unsigned long long* zero) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
    if (zero != nullptr && i == 0) *zero = 0ull;
}

// ---------------------------------------------------------------- sort ----
// The points of a Widom group sorted by cell in two counting passes: the
// generator's per-super-bin slots, a scan over super-bins, the points staged
// by super-bin, then one workgroup per super-bin counts and places them by
// cell in LDS.  Cells also get their slices (ceil(count / WG) work items each):
// offsets within the super-bin, then a scan over super-bins.
__device__ inline void scan_pair_1024(unsigned& p, unsigned& s, unsigned& tot_p, unsigned& tot_s) {
    // exclusive scan of (p, s) over the 1024 work-items; also the totals
    __shared__ unsigned w_p[SCAN_WG / 64], w_s[SCAN_WG / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned ip = p, is = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned tp = __shfl_up(ip, off, 64), ts = __shfl_up(is, off, 64);
        if (lane >= off) { ip += tp; is += ts; }
    }
    __syncthreads();
    if (lane == 63) { w_p[wave] = ip; w_s[wave] = is; }
    __syncthreads();
    unsigned bp = 0, bs = 0;
    tot_p = 0;
    tot_s = 0;
    for (int w = 0; w < SCAN_WG / 64; ++w) {
        if (w < wave) { bp += w_p[w]; bs += w_s[w]; }
        tot_p += w_p[w];
        tot_s += w_s[w];
    }
    p = bp + ip - p;
    s = bs + is - s;
}

// One workgroup: exclusive scan of in[0, n) into out[0, n], out[n] = the
// total, also stored at *total.
__global__ __launch_bounds__(SCAN_WG) void scan_excl_kernel(const unsigned* in, int n, unsigned* out,
                                                            unsigned* total) {
    unsigned run = 0;
    for (int base = 0; base < n; base += SCAN_WG) {
        const int k = base + threadIdx.x;
        unsigned p = k < n ? in[k] : 0u, s = 0u, tp, ts;
        scan_pair_1024(p, s, tp, ts);
        if (k < n) out[k] = run + p;
        run += tp;
    }
    if (threadIdx.x == 0) {
        out[n] = run;
        *total = run;
    }
}

// Every point to its super-bin's stretch of the staging array, (slot, cell);
// sb_off holds the offsets of the super-bins' counter copies, copy-minor.
__global__ void stage_points_kernel(const uint2* cell, long long n_pts, const unsigned* sb_off, int sb_shift,
                                    uint2* stage) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts) return;
    const uint2 c = cell[p];
    if (c.x == 0xffffffffu) return;
    stage[sb_off[((c.x & CELL_MASK) >> sb_shift) * SB_COPIES + (c.y >> 28)] + (c.y & SB_SLOT_MASK)] =
        make_uint2((unsigned)p, c.x);
}

// The second counting pass, one workgroup per super-bin: the super-bin's
// points counted per (cell, slab) in LDS, the cells' point offsets (the
// super-bin's points are a contiguous stretch of the sorted order) and, within
// the super-bin, their slice offsets (ceil(count / WG) slices per cell), then
// every point to its (cell, slab)'s stretch of `order`.
__global__ __launch_bounds__(SCAN_WG) void cell_sort_kernel(const uint2* stage, const unsigned* sb_off, int sb_shift,
                                                            int zstride, int n_cells, unsigned* count,
                                                            unsigned* point_off, unsigned* slice_loc,
                                                            unsigned* sb_slices, unsigned* order) {
    __shared__ unsigned s_n[SB_KEYS_MAX];
    const int tid = threadIdx.x;
    const int sb = blockIdx.x;
    const int c_lo = sb << sb_shift;
    const int nc = min(1 << sb_shift, n_cells - c_lo);
    const int nk = nc * zstride;
    const unsigned lo = sb_off[sb * SB_COPIES], hi = sb_off[(sb + 1) * SB_COPIES];
    for (int k = tid; k < nk; k += SCAN_WG) s_n[k] = 0u;
    __syncthreads();
    // four staged points per work-item in flight at once; key (cell, slab)
    for (unsigned i0 = lo + tid; i0 < hi; i0 += 4 * SCAN_WG) {
        unsigned cl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned i = i0 + j * SCAN_WG;
            cl[j] = i < hi ? stage[i].y : 0xffffffffu;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (cl[j] != 0xffffffffu)
                atomicAdd(&s_n[((cl[j] & CELL_MASK) - c_lo) * zstride + (cl[j] >> 24)], 1u);
    }
    __syncthreads();
    // each work-item a run of consecutive cells
    const int per = (nc + SCAN_WG - 1) / SCAN_WG;
    const int k0 = tid * per;
    unsigned p = 0u, s = 0u;
    for (int j = 0; j < per; ++j) {
        const int k = k0 + j;
        if (k < nc) {
            unsigned v = 0u;
            for (int z = 0; z < zstride; ++z) v += s_n[k * zstride + z];
            p += v;
            s += (v + WG - 1) / WG;
        }
    }
    unsigned tp, ts;
    scan_pair_1024(p, s, tp, ts);  // (its barriers also order the reads above before the writes below)
    p += lo;
    for (int j = 0; j < per; ++j) {
        const int k = k0 + j;
        if (k < nc) {
            const unsigned p0 = p;
            for (int z = 0; z < zstride; ++z) {
                const unsigned v = s_n[k * zstride + z];
                s_n[k * zstride + z] = p;  // from here on: the next free position of the slab
                p += v;
            }
            count[c_lo + k] = p - p0;
            point_off[c_lo + k] = p0;
            slice_loc[c_lo + k] = s;
            s += (p - p0 + WG - 1) / WG;
        }
    }
    if (tid == 0) sb_slices[sb] = ts;
    __syncthreads();
    for (unsigned i0 = lo + tid; i0 < hi; i0 += 4 * SCAN_WG) {
        uint2 e[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned i = i0 + j * SCAN_WG;
            e[j] = i < hi ? stage[i] : make_uint2(0u, 0xffffffffu);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (e[j].y != 0xffffffffu)
                order[atomicAdd(&s_n[((e[j].y & CELL_MASK) - c_lo) * zstride + (e[j].y >> 24)], 1u)] = e[j].x;
    }
}

__global__ void slice_items_kernel(const unsigned* count, const unsigned* slice_loc, const unsigned* sb_slice_off,
                                   int sb_shift, const SegArgs* segs, int n_seg, uint4* items, int n_cells) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int si = find_segment<SEG_BY_CELL>(segs, n_seg, c);
    unsigned ns = (count[c] + WG - 1) / WG, base = sb_slice_off[c >> sb_shift] + slice_loc[c];
    for (unsigned k = 0; k < ns; ++k) items[base + k] = make_uint4((unsigned)c, k, (unsigned)si, 0u);
}

// ---------------------------------------------------------------- pairs ----
// The erfc(alpha r) tables of a run in LDS: FP32 cubics (mixed path) or FP64
// quintics (FP64 path), each with 1/h of its segments.
struct ErfcTabs {
    float inv_h;
    const float4* tab;
    double inv_h64;
    const double2* tab64;
};

// The largest double below ERFC64_SEG (the FP64 twin of ERFC_XMAX).
constexpr double ERFC64_XMAX = (double)ERFC64_SEG - (double)ERFC64_SEG / 9007199254740992.0;

// erfc(alpha r) from the FP64 segment table, Horner in t = r/h - i.  The
// scaled distance is clamped below ERFC64_SEG: beyond r_sr (a masked term)
// the last segment gives a finite value, and so does r = NaN (r2 = 0, where
// the caller's 1/r is +inf).
__device__ inline double erfc_tab64(double r, const ErfcTabs& et) {
    const double x = fmin(r * et.inv_h64, ERFC64_XMAX);
    const int i = (int)x;
    const double t = x - (double)i;
    const double2 c01 = et.tab64[3 * i], c23 = et.tab64[3 * i + 1], c45 = et.tab64[3 * i + 2];
    return fma(fma(fma(fma(fma(c45.y, t, c45.x), t, c23.y), t, c23.x), t, c01.y), t, c01.x);
}

// 1/sqrt(r2) in FP64: v_rsq_f64 and two Newton steps y <- y (3/2 - r2 y^2 / 2)
// (the estimate's ~2^-22 relative error squares at each step, so the result
// is within a few ulp); r2 below the smallest normal gives +inf, as
// 1 / sqrt(0) does.
__device__ inline double rsq64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    const double h = 0.5 * r2;
    y = fma(y, fma(-(h * y), y, 0.5), y);
    y = fma(y, fma(-(h * y), y, 0.5), y);
    return r2 >= 2.2250738585072014e-308 ? y : INFINITY;
}

// One candidate pair: LJ into plj, real-space Ewald into psr.
// FP32 pairs: the masked terms accumulate by one FMA each (eps4 * w and
// q / r * erfc), the table offset is v_fract of the scaled distance, and the
// candidate's charge arrives in FP32 (qf).  FP64 pairs: 1/r from rsq64,
// 1/r^2 as its square, erfc(alpha r) from the FP64 table.
template <int ESM, typename Real>
__device__ inline void pair_term(double r2, double q, float qf, const LJP<Real>& p, bool is_lj, double r2lj,
                                 double r2sr, const ErfcTabs& et, bool& ovl, Real& plj, Real& psr) {
    const float inv_h = et.inv_h;
    const float4* tab = et.tab;
    const bool in = r2 < r2lj;
    ovl |= in & (r2 < R2_OVERLAP) & is_lj;
    if constexpr (ESM != 0 && sizeof(Real) == 4) {
        // one transcendental per pair: 1/r from v_rsq, 1/r^2 as its square
        const float rf2 = (float)r2;
        const float rinv = __builtin_amdgcn_rsqf(rf2);
        const float s = p.sig2 * (rinv * rinv);
        const float s3 = s * s * s;
        plj = fmaf(p.eps4, in ? fmaf(s3, s3, -s3) : 0.0f, plj);
        const bool ins = ((ESM == 1) ? in : (r2 < r2sr)) & (qf != 0.0f);
        const float x = fminf(rf2 * rinv * inv_h, ERFC_XMAX);
        const int i = (int)x;
        const float t = __builtin_amdgcn_fractf(x);
        const float4 c = tab[i];
        const float ec = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
        psr = fmaf(ins ? qf * rinv : 0.0f, ec, psr);
    } else if constexpr (sizeof(Real) == 4) {
        const float s = p.sig2 * __builtin_amdgcn_rcpf((float)r2);
        const float s3 = s * s * s;
        plj = fmaf(p.eps4, in ? fmaf(s3, s3, -s3) : 0.0f, plj);
    } else {
        const double y = rsq64(r2);
        const double s = p.sig2 * (y * y);
        const double s3 = s * s * s;
        plj += in ? p.eps4 * s3 * (s3 - 1.0) : 0.0;
        if constexpr (ESM != 0) {
            const bool ins = ((ESM == 1) ? in : (r2 < r2sr)) & (q != 0.0);
            psr += ins ? q * y * erfc_tab64(r2 * y, et) : 0.0;
        }
    }
}

// pair_term with the sums of a gradient run, for a candidate of the framework
// type the caller is summing: sw += u^2 - u and su += u over pairs inside the
// LJ cutoff (u = (sigma_ts^2 / r^2)^3), sq += erfc(alpha r) / r over pairs
// inside the real-space cutoff, all in Real.  d LJ / d eps_ts = 4 (u^2 - u) and
// d LJ / d sigma_ts = (4 eps_ts / sigma_ts)(12 u^2 - 6 u) = csig (12 sw + 6 su),
// so the caller forms both per type from sw and su.  The energy terms are
// pair_term's, operation for operation: a gradient run's energies equal a
// plain run's bit for bit.
template <int ESM, typename Real>
__device__ inline void pair_term_grad(double r2, double q, float qf, const LJP<Real>& p, bool is_lj, double r2lj,
                                      double r2sr, const ErfcTabs& et, bool& ovl, Real& plj, Real& psr, Real& sw,
                                      Real& su, Real& sq) {
    const bool in = r2 < r2lj;
    ovl |= in & (r2 < R2_OVERLAP) & is_lj;
    if constexpr (sizeof(Real) == 4) {
        const float rf2 = (float)r2;
        float rinv = 0.0f, s;
        if constexpr (ESM != 0) {
            rinv = __builtin_amdgcn_rsqf(rf2);
            s = p.sig2 * (rinv * rinv);
        } else {
            s = p.sig2 * __builtin_amdgcn_rcpf(rf2);
        }
        const float s3 = s * s * s;
        const float wm = in ? fmaf(s3, s3, -s3) : 0.0f;
        plj = fmaf(p.eps4, wm, plj);
        sw += wm;
        su += in ? s3 : 0.0f;
        if constexpr (ESM != 0) {
            const bool inr = (ESM == 1) ? in : (r2 < r2sr);
            const float x = fminf(rf2 * rinv * et.inv_h, ERFC_XMAX);
            const int i = (int)x;
            const float t = __builtin_amdgcn_fractf(x);
            const float4 c = et.tab[i];
            const float ec = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
            psr = fmaf((inr & (qf != 0.0f)) ? qf * rinv : 0.0f, ec, psr);
            sq += inr ? ec * rinv : 0.0f;
        }
    } else {
        const double y = rsq64(r2);
        const double s = p.sig2 * (y * y);
        const double s3 = s * s * s;
        plj += in ? p.eps4 * s3 * (s3 - 1.0) : 0.0;
        sw += in ? s3 * (s3 - 1.0) : 0.0;
        su += in ? s3 : 0.0;
        if constexpr (ESM != 0) {
            const bool inr = (ESM == 1) ? in : (r2 < r2sr);
            const double ec = erfc_tab64(r2 * y, et);
            psr += (inr & (q != 0.0)) ? q * y * ec : 0.0;
            sq += inr ? ec * y : 0.0;
        }
    }
}

// FP32 kernels pack a candidate's type offset (high word) and FP32 charge
// (low word) into the fourth double of its record, so one 32-byte LDS record
// carries the candidate (raw bits: no arithmetic touches the packed value).
__device__ inline double pack_tq(int t, float q) {
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)t << 32) | (unsigned)__float_as_int(q)));
}
__device__ inline int unpack_t(double w) { return (int)(__double_as_longlong(w) >> 32); }
__device__ inline float unpack_q(double w) { return __int_as_float((int)__double_as_longlong(w)); }

struct PairLds {
    double4* c;     // candidate records: Cartesian xyz; charge (FP64 kernels) or packed type and charge
    int* t;         // candidate type * nsite
    float* qf;      // candidate charge in FP32
    int* sel;       // selected tiles of the current batch
    unsigned* ends; // gradient runs: per wave's tile, the cumulative kept candidates of types 0..k,
                    // 7 bits per type (two words)
    const float4* erfc;    // FP32 kernels
    const double2* erfc64;  // FP64 kernels
    int* wcnt;      // per-wave candidate counts
    double* sph;    // per-wave bounding-box and radius partials of the workgroup's sphere
};

// Per-type derivative sums of a gradient run's point, FP64, for NT framework
// types: g[k] = sum (u^2 - u) (d LJ / d eps_ts = 4 of it), g[NT + k] =
// d LJ / d sigma_ts, g[2 NT + k] = sum erfc(alpha r) / r.  A run of candidates
// of one type accumulates in Real (pair_term_grad's sw, su, sq) and is folded
// in here once.  The image-list path folds with the type known at compile time
// (add), the all-atoms path with the type in a register (add_dyn, every entry
// updated: a conditional update of one entry would be folded into a
// dynamically indexed array held in scratch memory).
template <typename Real, int NT>
struct TypeSums {
    double g[3 * NT];
    // the all-atoms path: the type (times nsite) being summed and its sums
    int cur = -1;
    Real sw = (Real)0, su = (Real)0, sq = (Real)0, csig = (Real)0;
    __device__ TypeSums() {
#pragma unroll
        for (int k = 0; k < 3 * NT; ++k) g[k] = 0.0;
    }
    __device__ __attribute__((always_inline)) inline void add(int k, Real w, Real u, Real q, Real cs, bool is_lj) {
        g[k] += is_lj ? (double)w : 0.0;
        g[NT + k] += is_lj ? (double)cs * (12.0 * (double)w + 6.0 * (double)u) : 0.0;
        g[2 * NT + k] += (double)q;
    }
    __device__ __attribute__((always_inline)) inline void add_dyn(int t, bool is_lj) {
        const double e = is_lj ? (double)sw : 0.0;
        const double d = is_lj ? (double)csig * (12.0 * (double)sw + 6.0 * (double)su) : 0.0;
        const double c = (double)sq;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            g[k] += (k == t) ? e : 0.0;
            g[NT + k] += (k == t) ? d : 0.0;
            g[2 * NT + k] += (k == t) ? c : 0.0;
        }
        sw = su = sq = (Real)0;
    }
    // all-atoms path: a candidate of type tt (times nsite) with d/d sigma factor cs
    __device__ __attribute__((always_inline)) inline void type(int tt, int nsite, Real cs, bool is_lj) {
        if (tt != cur) {
            if (cur >= 0) add_dyn(cur / nsite, is_lj);
            cur = tt;
            csig = cs;
        }
    }
    __device__ __attribute__((always_inline)) inline void flush(int nsite, bool is_lj) {
        if (cur >= 0) add_dyn(cur / nsite, is_lj);
    }
};

// The pair loops of one workgroup's chunk.  Fast path: bound the workgroup's
// points by a sphere, scan the image list, keep the images within r_cmax of
// the sphere by a wave-ballot compaction into LDS, evaluate every lane's point
// against them (an image outside a point's cutoffs adds exact zeros, so the
// sums do not depend on the sphere).
// Slow path: every replicated atom, nearest image per pair.  NT > 0 (a
// gradient run of at most NT framework types) adds the per-type derivative
// sums: on the fast path the images of a tile come in type order (the image
// list is sorted by species inside each spatial bin, and a tile never spans
// two bins), so each tile's kept candidates are runs of one type each, counted
// per type at compaction; a run is summed with its type's parameters held in
// registers and folded into the type's FP64 sums once.
template <int ESM, typename Real, int NT>
__device__ __attribute__((always_inline)) inline void chunk_pairs(const RunArgs& a, bool valid, double px, double py, double pz,
                                   int site, const LJP<Real>* s_lj, const Real* s_gc,
                                   PairLds L, double& acc_lj, double& acc_sr, bool& ovl, unsigned long long& n_cand,
                                   TypeSums<Real, (NT > 0 ? NT : 1)>& ts) {
    constexpr bool GRAD = NT > 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double f11 = a.f11, f12 = a.f12, f13 = a.f13, f22 = a.f22, f23 = a.f23, f33 = a.f33;
    const double r2lj = a.r2_lj, r2sr = a.r2_sr;
    const ErfcTabs et{(float)a.erfc_inv_h, L.erfc, a.erfc64_inv_h, L.erfc64};
    const bool is_lj = (a.site_lj_mask >> site) & 1u;
    const int nsite = a.nsite;
    // a wave with no valid point skips the pair loops; it still takes part in
    // the tile scans and barriers
    const bool wave_live = __ballot(valid) != 0ull;
    if (a.fast) {
        // The point and the chunk centre in Cartesian coordinates (FP64).  The
        // images are stored in Cartesian coordinates, so a pair's r^2 depends
        // on the point and the image alone, not on the chunk it was found from.
        const double prx = f11 * px + f12 * py + f13 * pz;
        const double pry = f22 * py + f23 * pz;
        const double prz = f33 * pz;
        // The workgroup's sphere: the centre of its valid points' bounding box
        // and their largest distance from it (inside the chunk's sphere, and
        // smaller where the points leave the chunk's corners empty).
        double bx[6] = {valid ? prx : INFINITY,  valid ? pry : INFINITY,  valid ? prz : INFINITY,
                        valid ? -prx : INFINITY, valid ? -pry : INFINITY, valid ? -prz : INFINITY};
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
            for (int i = 0; i < 6; ++i) bx[i] = fmin(bx[i], __shfl_xor(bx[i], off, 64));
        if (lane == 0)
#pragma unroll
            for (int i = 0; i < 6; ++i) L.sph[wave * 8 + i] = bx[i];
        __syncthreads();
#pragma unroll
        for (int w = 0; w < WG / 64; ++w)
#pragma unroll
            for (int i = 0; i < 6; ++i) bx[i] = fmin(bx[i], L.sph[w * 8 + i]);
        const double ccx = 0.5 * (bx[0] - bx[3]), ccy = 0.5 * (bx[1] - bx[4]), ccz = 0.5 * (bx[2] - bx[5]);
        double d2 = 0.0;
        if (valid) {
            const double ex = prx - ccx, ey = pry - ccy, ez = prz - ccz;
            d2 = ex * ex + ey * ey + ez * ez;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off, 64));
        if (lane == 0) L.sph[wave * 8 + 6] = d2;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < WG / 64; ++w) d2 = fmax(d2, L.sph[w * 8 + 6]);
        const double Rc = a.r_cmax + sqrt(d2) + 1e-3;
        const double R2 = Rc * Rc;
        const int n_tiles = a.n_tiles;
        for (int tb = 0; tb < n_tiles; tb += WG) {
            // the tiles of this batch whose bounding sphere meets the chunk's
            bool hit = false;
            const int ti = tb + tid;
            if (ti < n_tiles) {
                const double4 sp = a.tile_sph[ti];
                const double sx = sp.x - ccx, sy = sp.y - ccy, sz = sp.z - ccz;
                const double rr = Rc + sp.w;
                hit = sx * sx + sy * sy + sz * sz < rr * rr;
            }
            const unsigned long long hm = __ballot(hit);
            const int hb = __popcll(hm & ((1ull << lane) - 1ull));
            __syncthreads();  // the previous batch's tile list and candidates are consumed
            if (lane == 0) L.wcnt[wave] = __popcll(hm);
            __syncthreads();
            int hoff = 0, n_hit = 0;
#pragma unroll
            for (int w = 0; w < WG / 64; ++w) {
                const int c = L.wcnt[w];
                hoff += w < wave ? c : 0;
                n_hit += c;
            }
            if (hit) L.sel[hoff + hb] = ti;
            __syncthreads();
            // rounds of WG / 64 selected tiles, one per wave, in tile order
            for (int k = 0; k < n_hit; k += WG / 64) {
                bool keep = false;
                double4 rec = make_double4(0.0, 0.0, 0.0, 0.0);
                int ty = 0;
                if (k + wave < n_hit) {
                    const int2 rg = a.tile_rng[L.sel[k + wave]];
                    if (lane < rg.y) {
                        rec = a.img[rg.x + lane];
                        const double cx = rec.x - ccx, cy = rec.y - ccy, cz = rec.z - ccz;
                        keep = cx * cx + cy * cy + cz * cz < R2;
                        ty = a.img_type[rg.x + lane] * nsite;
                    }
                }
                const unsigned long long m = __ballot(keep);
                const int before = __popcll(m & ((1ull << lane) - 1ull));
                int tc[NT > 0 ? NT : 1];
                if constexpr (GRAD) {
#pragma unroll
                    for (int k = 0; k < NT; ++k) tc[k] = __popcll(__ballot(keep && ty == k * nsite));
                }
                __syncthreads();  // the previous round's candidates (and the tile counts) are consumed
                if (lane == 0) {
                    L.wcnt[wave] = __popcll(m);
                    if constexpr (GRAD) {
                        // cumulative counts, 7 bits per type (a tile holds at most 64)
                        unsigned long long ends = 0ull;
                        int cum = 0;
#pragma unroll
                        for (int k = 0; k < NT; ++k) {
                            cum += tc[k];
                            ends |= (unsigned long long)cum << (7 * k);
                        }
                        L.ends[2 * wave] = (unsigned)ends;
                        L.ends[2 * wave + 1] = (unsigned)(ends >> 32);
                    }
                }
                __syncthreads();
                int woff[WG / 64 + 1];
                woff[0] = 0;
#pragma unroll
                for (int w = 0; w < WG / 64; ++w) woff[w + 1] = woff[w] + L.wcnt[w];
                if (keep) {
                    if constexpr (sizeof(Real) == 4) {
                        L.c[woff[wave] + before] = make_double4(rec.x, rec.y, rec.z, pack_tq(ty, (float)rec.w));
                    } else {
                        L.c[woff[wave] + before] = rec;
                        L.t[woff[wave] + before] = ty;
                    }
                }
                __syncthreads();
                n_cand += (unsigned long long)woff[WG / 64];
                if (wave_live) {
                    if constexpr (GRAD) {
                        // tile by tile (the energies' FP32 sums fold per tile, as below),
                        // type run by type run inside each tile; the runs' ends come
                        // from one scalar word per tile
                        int j = 0;
                        for (int gi = 0; gi < WG / 64; ++gi) {
                            Real plj = (Real)0, psr = (Real)0;
                            const int j0 = j;
                            const unsigned long long ends =
                                (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(L.ends[2 * gi]) |
                                (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(L.ends[2 * gi + 1])
                                    << 32;
#pragma unroll
                            for (int k = 0; k < NT; ++k) {
                                const int je = j0 + (int)((ends >> (7 * k)) & 127ull);
                                if (je == j) continue;
                                const LJP<Real> p = s_lj[k * nsite + site];
                                const Real cs = s_gc[k * nsite + site];
                                Real sw = (Real)0, su = (Real)0, sq = (Real)0;
#pragma unroll 2
                                for (; j < je; ++j) {
                                    const double4 c = L.c[j];
                                    const double ddx = c.x - prx, ddy = c.y - pry, ddz = c.z - prz;
                                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                    const float qj = sizeof(Real) == 4 ? unpack_q(c.w) : 0.0f;
                                    pair_term_grad<ESM, Real>(r2, c.w, qj, p, is_lj, r2lj, r2sr, et, ovl, plj, psr, sw,
                                                              su, sq);
                                }
                                ts.add(k, sw, su, sq, cs, is_lj);
                            }
                            acc_lj += (double)plj;
                            acc_sr += (double)psr;
                        }
                    } else {
                        // the candidates of each tile accumulate in Real, flushed to FP64
#pragma unroll
                        for (int gi = 0; gi < WG / 64; ++gi) {
                            Real plj = (Real)0, psr = (Real)0;
                            // two candidates per iteration share the LDS address registers (FP32 energies)
                            constexpr int PAIR_UNROLL = sizeof(Real) == 4 ? 2 : 1;
#pragma unroll PAIR_UNROLL
                            for (int j = woff[gi]; j < woff[gi + 1]; ++j) {
                                const double4 c = L.c[j];
                                const double ddx = c.x - prx, ddy = c.y - pry, ddz = c.z - prz;
                                const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                const bool f32 = sizeof(Real) == 4;
                                const int tj = f32 ? unpack_t(c.w) : L.t[j];
                                const float qj = f32 ? unpack_q(c.w) : 0.0f;  // the FP64 terms take c.w
                                pair_term<ESM, Real>(r2, c.w, qj, s_lj[tj + site], is_lj, r2lj, r2sr, et, ovl, plj,
                                                     psr);
                            }
                            acc_lj += (double)plj;
                            acc_sr += (double)psr;
                        }
                    }
                }
            }
        }
    } else {
        const int n_atoms = a.n_atoms;
        for (int base = 0; base < n_atoms; base += TILE) {
            const int nt = min(TILE, n_atoms - base);
            __syncthreads();
            if (tid < nt) {
                L.c[tid] = a.atoms[base + tid];
                L.t[tid] = a.species[base + tid] * nsite;
                L.qf[tid] = (float)a.atoms[base + tid].w;
            }
            __syncthreads();
            n_cand += (unsigned long long)nt;
            if (wave_live) {
                for (int gi = 0; gi < nt; gi += 64) {
                    Real plj = (Real)0, psr = (Real)0;
                    const int ge = min(nt, gi + 64);
                    for (int j = gi; j < ge; ++j) {
                        const double4 at = L.c[j];
                        double ex = at.x - px, ey = at.y - py, ez = at.z - pz;
                        ex -= rint(ex);
                        ey -= rint(ey);
                        ez -= rint(ez);
                        const double cx = f11 * ex + f12 * ey + f13 * ez;
                        const double cy = f22 * ey + f23 * ez;
                        const double cz = f33 * ez;
                        const double r2 = cx * cx + cy * cy + cz * cz;
                        if constexpr (GRAD) {
                            const int tt = __builtin_amdgcn_readfirstlane(L.t[j]);
                            ts.type(tt, nsite, s_gc[tt + site], is_lj);
                            pair_term_grad<ESM, Real>(r2, at.w, L.qf[j], s_lj[tt + site], is_lj, r2lj, r2sr, et,
                                                      ovl, plj, psr, ts.sw, ts.su, ts.sq);
                        } else {
                            pair_term<ESM, Real>(r2, at.w, L.qf[j], s_lj[L.t[j] + site], is_lj, r2lj, r2sr, et, ovl,
                                                 plj, psr);
                        }
                    }
                    acc_lj += (double)plj;
                    acc_sr += (double)psr;
                    if constexpr (GRAD) ts.flush(nsite, is_lj);  // FP64 per 64 candidates, as the energies
                }
            }
        }
    }
    if constexpr (GRAD) ts.flush(nsite, is_lj);
}

// LDS: TILE candidate records, types and FP32 charges, the selected-tile list,
// the erfc table (FP32 cubics or FP64 quintics), then the LJ table of the run
// (and for a gradient run its d/d sigma factors).
template <typename Real>
constexpr int pair_lds_fixed() {
    return TILE * (int)(sizeof(double4) + sizeof(int) + sizeof(float) + sizeof(int)) +
           (sizeof(Real) == 4 ? ERFC_SEG * (int)sizeof(float4) : 3 * ERFC64_SEG * (int)sizeof(double2));
}

// NT = 0: energies.  NT > 0: a gradient run of at most NT framework types.
template <typename Real, int NT>
__global__ __launch_bounds__(WG) void point_pair_kernel(GroupArgs g) {
    constexpr bool GRAD = NT > 0;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_wcnt[WG / 64];
    __shared__ unsigned s_ends[(WG / 64) * 2];
    __shared__ double s_sph[(WG / 64) * 8];
    PairLds L;
    L.c = reinterpret_cast<double4*>(smem);
    L.t = reinterpret_cast<int*>(smem + TILE * sizeof(double4));
    L.qf = reinterpret_cast<float*>(smem + TILE * (sizeof(double4) + sizeof(int)));
    L.sel = reinterpret_cast<int*>(smem + TILE * (sizeof(double4) + sizeof(int) + sizeof(float)));
    float4* s_erfc = reinterpret_cast<float4*>(smem + TILE * (sizeof(double4) + 2 * sizeof(int) + sizeof(float)));
    L.erfc = s_erfc;
    double2* s_erfc64 = reinterpret_cast<double2*>(s_erfc);
    L.erfc64 = s_erfc64;
    L.wcnt = s_wcnt;
    L.ends = s_ends;
    L.sph = s_sph;
    LJP<Real>* s_lj = reinterpret_cast<LJP<Real>*>(smem + pair_lds_fixed<Real>());

    const int tid = threadIdx.x;
    int si;
    uint4 it = make_uint4(0u, 0u, 0u, 0u);
    if (g.mode == MODE_WIDOM) {
        if (blockIdx.x >= *g.n_items) return;  // the whole workgroup leaves
        it = g.items[blockIdx.x];
        si = (int)it.z;
    } else {
        si = find_segment<SEG_BY_WG>(g.segs, g.n_seg, blockIdx.x);
    }
    const SegArgs sg = g.segs[si];
    const RunArgs& a = g.runs[sg.run];
    Real* s_gc = reinterpret_cast<Real*>(s_lj + a.n_types * a.nsite);

    bool valid;
    long long pid;
    double px, py, pz;
    int site;
    if (g.mode == MODE_WIDOM) {
        const unsigned k = g.point_off[it.x] + it.y * WG + tid;
        valid = k < g.point_off[it.x + 1];
        pid = valid ? (long long)g.order[k] : sg.slot_base;
        const double4 p = valid ? g.pos[pid] : make_double4(0.0, 0.0, 0.0, 0.0);
        px = p.x; py = p.y; pz = p.z;
        site = (int)((pid - sg.slot_base) % a.nsite);
    } else {
        const int lb = (int)blockIdx.x - sg.wg_base;
        const int nb0 = a.nbrick[0], nb1 = a.nbrick[1];
        const int bi = lb % nb0, bj = (lb / nb0) % nb1, bk = lb / (nb0 * nb1);
        const int i0 = bi * a.brick[0], j0 = bj * a.brick[1], k0 = bk * a.brick[2];
        const int ti = tid % a.brick[0], tj = (tid / a.brick[0]) % a.brick[1], tk = tid / (a.brick[0] * a.brick[1]);
        const int i = i0 + ti, j = j0 + tj, k = k0 + tk;
        const int n1 = a.grid_n[0], n2 = a.grid_n[1], n3 = a.grid_n[2];
        valid = i < n1 && j < n2 && k < n3;
        pid = sg.slot_base + (valid ? (long long)i + (long long)n1 * ((long long)j + (long long)n2 * k) : 0);
        px = grid_coord(i, n1, a.repd[0]);
        py = grid_coord(j, n2, a.repd[1]);
        pz = grid_coord(k, n3, a.repd[2]);
        site = 0;
    }
    const int n_tab = a.n_types * a.nsite;
    const LJP<Real>* gtab = reinterpret_cast<const LJP<Real>*>(a.ljtab);
    for (int t = tid; t < n_tab; t += WG) s_lj[t] = gtab[t];
    if constexpr (GRAD) {
        const Real* gc = reinterpret_cast<const Real*>(a.gtab);
        for (int t = tid; t < n_tab; t += WG) s_gc[t] = gc[t];
    }
    if constexpr (sizeof(Real) == 4)
        if (a.esm != 0)
            for (int t = tid; t < ERFC_SEG; t += WG) s_erfc[t] = a.erfc_tab[t];
    if constexpr (sizeof(Real) == 8)
        if (a.esm != 0)
            for (int t = tid; t < 3 * ERFC64_SEG; t += WG) s_erfc64[t] = a.erfc64_tab[t];

    double acc_lj = 0.0, acc_sr = 0.0;
    bool ovl = false;
    unsigned long long n_cand = 0;
    TypeSums<Real, (NT > 0 ? NT : 1)> ts;
    if (a.esm == 0)
        chunk_pairs<0, Real, NT>(a, valid, px, py, pz, site, s_lj, s_gc, L, acc_lj, acc_sr, ovl,
                                 n_cand, ts);
    else if (a.esm == 1)
        chunk_pairs<1, Real, NT>(a, valid, px, py, pz, site, s_lj, s_gc, L, acc_lj, acc_sr, ovl,
                                 n_cand, ts);
    else
        chunk_pairs<2, Real, NT>(a, valid, px, py, pz, site, s_lj, s_gc, L, acc_lj, acc_sr, ovl,
                                 n_cand, ts);
    if (valid) {
        g.out_lj[pid] = ovl ? INFINITY : acc_lj;
        g.out_sr[pid] = acc_sr;
        if constexpr (GRAD) {
            // (types NT .. GT - 1 of the fixed stride are never read)
            double* o = g.out_g + pid;
            const long long gs = g.g_stride;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                o[k * gs] = ts.g[k];
                o[(GT + k) * gs] = ts.g[NT + k];
                o[(2 * GT + k) * gs] = ts.g[2 * NT + k];
            }
        }
    }
    if (g.cand_pairs != nullptr) {
        const unsigned long long m = __ballot(valid);
        if ((tid & 63) == 0 && m != 0ull)
            atomicAdd(g.cand_pairs, (unsigned long long)__popcll(m) * n_cand);
    }
}

// -------------------------------------------------------------- Fourier ----
__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Points per work-item of the Fourier kernel: the points of a work-item share
// every coefficient load and the loop control, and their Horner chains
// interleave.
constexpr int FOURIER_PTS = 2;
constexpr int FOURIER_WG = WG / FOURIER_PTS;

struct FourierPoint {
    long long pid;
    bool valid, charged;
    double2 Ea, Eb;  // z_a, z_b
    double zr, zi;   // z_c
};

__device__ inline FourierPoint fourier_point(const GroupArgs& g, const SegArgs& sg, const RunArgs& a,
                                             long long pid) {
    FourierPoint P;
    P.pid = pid;
    const long long lp = pid - sg.slot_base;
    bool valid = lp < sg.n_slots;
    double x = 0.0, y = 0.0, z = 0.0;
    int site = 0;
    if (g.mode == MODE_WIDOM) {
        const double4 p = valid ? g.pos[pid] : make_double4(0.0, 0.0, 0.0, 0.0);
        valid = valid && p.w != 0.0;
        x = p.x; y = p.y; z = p.z;
        site = (int)(lp % a.nsite);
    } else if (valid) {
        const int n1 = a.grid_n[0], n2 = a.grid_n[1], n3 = a.grid_n[2];
        const long long i = lp % n1, j = (lp / n1) % n2, k = lp / ((long long)n1 * n2);
        x = grid_coord((double)i, n1, a.repd[0]);
        y = grid_coord((double)j, n2, a.repd[1]);
        z = grid_coord((double)k, n3, a.repd[2]);
    }
    P.valid = valid;
    P.charged = valid && ((a.site_q_mask >> site) & 1u);
    double t0 = x * a.repd[0], t1 = y * a.repd[1], t2 = z * a.repd[2];
    t0 -= rint(t0);
    t1 -= rint(t1);
    t2 -= rint(t2);
    double s, c;
    sincospi(2.0 * t0, &s, &c);
    P.Ea = make_double2(c, -s);
    sincospi(2.0 * t1, &s, &c);
    P.Eb = make_double2(c, -s);
    sincospi(2.0 * t2, &s, &c);
    P.zr = c;
    P.zi = -s;
    return P;
}

// Fourier part per point: sum_m Re(A_m e^{-i 2 pi m.u}), u = R o x (unit-cell
// fractional), FP64.  With z_a = e^{-i 2 pi u_a}: for every (ma, mb) row the
// sum over the row's kept mc range is evaluated as two polynomials by Horner,
// sum_{mc >= 0} A_mc z_c^mc and z_c^* sum_{j >= 1} A_{-j} (z_c^*)^{j-1}, and
// multiplied by z_a^ma z_b^mb, carried along the rows by recurrence.  The
// coefficients are the same for every lane: they are read by scalar loads
// (through the scalar cache, into SGPR operands of the FMAs) rather than
// broadcast from LDS, which would move 64 copies of each one.  Point slots of
// a segment start at a multiple of WG, so a workgroup lies in one segment.
__global__ __launch_bounds__(FOURIER_WG) void point_fourier_kernel(GroupArgs g) {
    const int tid = threadIdx.x;
    const long long p0 = (long long)blockIdx.x * WG;  // WG slots per workgroup, FOURIER_PTS per work-item
    const SegArgs sg = g.segs[find_segment<SEG_BY_SLOT>(g.segs, g.n_seg, p0)];
    const RunArgs& a = g.runs[sg.run];
    if (a.esm == 0) return;  // uniform: the whole workgroup leaves
    FourierPoint P[FOURIER_PTS];
    double nzr[FOURIER_PTS], nzi[FOURIER_PTS];
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k) {
        P[k] = fourier_point(g, sg, a, p0 + tid + k * FOURIER_WG);
        // opaque to the optimiser, which would fold -q * -z back to q * z
        nzr[k] = -P[k].zr;
        nzi[k] = -P[k].zi;
        asm volatile("" : "+v"(nzr[k]));
        asm volatile("" : "+v"(nzi[k]));
    }
    // wave-uniform: scalar loop control (the loads are not provably uniform)
    const int Ma = __builtin_amdgcn_readfirstlane(a.Ma), Mb = __builtin_amdgcn_readfirstlane(a.Mb),
              Mc = __builtin_amdgcn_readfirstlane(a.Mc);
    // per m_a: T = sum_mb z_b^(mb + Mb) R(ma, mb) by Horner in z_b over the rows
    // (mb from Mb down to -Mb), then acc += Re(z_a^ma z_b^-Mb T)
    double2 Eb_neg[FOURIER_PTS], Pa[FOURIER_PTS];  // z_b^-Mb, z_a^ma
    double acc[FOURIER_PTS];
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k) {
        Eb_neg[k] = make_double2(1.0, 0.0);
        const double2 Eb_c = make_double2(P[k].Eb.x, -P[k].Eb.y);
        for (int q = 0; q < Mb; ++q) Eb_neg[k] = cmul(Eb_neg[k], Eb_c);
        Pa[k] = make_double2(1.0, 0.0);
        acc[k] = 0.0;
    }

    const int rowlen = 2 * Mc + 1, nb = 2 * Mb + 1;
    // (plain scalars through these pointers: the vector types have no copy
    // from that address space)
    const gconst_ptr<double> Abox = uniform_ptr(reinterpret_cast<const double*>(a.A));
    const gconst_ptr<int> krows = uniform_ptr(reinterpret_cast<const int*>(a.krows));
    for (int ma = 0; ma <= Ma; ++ma) {
        double Tr[FOURIER_PTS], Ti[FOURIER_PTS];
#pragma unroll
        for (int k = 0; k < FOURIER_PTS; ++k) Tr[k] = Ti[k] = 0.0;
        for (int jb = nb - 1; jb >= 0; --jb) {
            const int r = ma * nb + jb;
            const int2 lh = make_int2(krows[2 * r], krows[2 * r + 1]);
            double Rr[FOURIER_PTS], Ri[FOURIER_PTS];  // the row's sum_mc A z_c^mc
#pragma unroll
            for (int k = 0; k < FOURIER_PTS; ++k) Rr[k] = Ri[k] = 0.0;
            if (lh.x <= lh.y) {
                const gconst_ptr<double> row = Abox + 2 * ((long long)r * rowlen + Mc);  // mc in [-Mc, Mc]
                // Every inner FMA carries a negated operand (-q * -z where the sign
                // is not wanted), which keeps it in the three-operand form that
                // takes the scalar coefficient directly; the products are exact
                // negations of each other, so the sums are unchanged.
                double qr[FOURIER_PTS], qi[FOURIER_PTS];  // sum_{mc >= 0} A_mc z^mc
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) qr[k] = qi[k] = 0.0;
#pragma unroll 2
                for (int mc = lh.y; mc >= 0; --mc) {
                    const double Ar = row[2 * mc], Ai = row[2 * mc + 1];
#pragma unroll
                    for (int k = 0; k < FOURIER_PTS; ++k) {
                        const double nr = fma(qr[k], P[k].zr, fma(-qi[k], P[k].zi, Ar));
                        qi[k] = fma(qr[k], P[k].zi, fma(-qi[k], nzr[k], Ai));
                        qr[k] = nr;
                    }
                }
                double nr_[FOURIER_PTS], ni_[FOURIER_PTS];  // sum_{j >= 1} A_{-j} (z*)^{j-1}
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) nr_[k] = ni_[k] = 0.0;
#pragma unroll 2
                for (int j = -lh.x; j >= 1; --j) {
                    const double Ar = row[-2 * j], Ai = row[-2 * j + 1];
#pragma unroll
                    for (int k = 0; k < FOURIER_PTS; ++k) {
                        const double tr = fma(nr_[k], P[k].zr, fma(-ni_[k], nzi[k], Ar));
                        ni_[k] = fma(ni_[k], P[k].zr, fma(-nr_[k], P[k].zi, Ai));
                        nr_[k] = tr;
                    }
                }
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) {
                    // R = q + z* n
                    const double zr = P[k].zr, zi = P[k].zi;
                    Rr[k] = qr[k] + (zr * nr_[k] + zi * ni_[k]);
                    Ri[k] = qi[k] + (zr * ni_[k] - zi * nr_[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < FOURIER_PTS; ++k) {
                // T = T z_b + R
                const double tr = fma(Tr[k], P[k].Eb.x, fma(-Ti[k], P[k].Eb.y, Rr[k]));
                Ti[k] = fma(Tr[k], P[k].Eb.y, fma(Ti[k], P[k].Eb.x, Ri[k]));
                Tr[k] = tr;
            }
        }
#pragma unroll
        for (int k = 0; k < FOURIER_PTS; ++k) {
            const double2 Q = cmul(Pa[k], Eb_neg[k]);
            acc[k] += Q.x * Tr[k] - Q.y * Ti[k];
            Pa[k] = cmul(Pa[k], P[k].Ea);
        }
    }
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k)
        if (P[k].valid) g.out_lr[P[k].pid] = P[k].charged ? acc[k] : 0.0;
}

// Fourier part of an energy grid, separably.  On a grid the phase of point
// (i, j, k) factorises, e^{-2 pi i m.u} = z_a(i)^ma z_b(j)^mb z_c(k)^mc, so
// sum_m A_m z^m is three nested one-axis sums: B(ma, mb, k) = sum_mc A z_c^mc,
// C(ma, j, k) = sum_mb B z_b^mb, lr(i, j, k) = Re sum_ma C z_a^ma, with work
// O(n (Ma + 1)) for n points instead of O(n x box).  FP64 throughout; the
// phases are computed exactly as in point_fourier_kernel.
struct GridFourierArgs {
    const double2* A;  // the run's box: ((ma (2 Mb + 1)) + mb + Mb) (2 Mc + 1) + mc + Mc
    int Ma, Mb, Mc;
    int n[3];
    double repd[3];
    long long slot_base;
    double* out_lr;
    double2* z;        // phases: n[0] of z_a, then n[1] of z_b, then n[2] of z_c
    double2* B;        // (Ma + 1)(2 Mb + 1) x n[2]
    double2* C;        // (Ma + 1) x n[1] x n[2]
};

__device__ inline double2 cfma(double2 h, double2 z, double2 a) {  // h z + a
    return make_double2(fma(h.x, z.x, fma(-h.y, z.y, a.x)), fma(h.x, z.y, fma(h.y, z.x, a.y)));
}

__global__ void grid_phase_kernel(GridFourierArgs f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0;
    if (i >= f.n[0]) { i -= f.n[0]; d = 1; }
    if (d == 1 && i >= f.n[1]) { i -= f.n[1]; d = 2; }
    if (d == 2 && i >= f.n[2]) return;
    double t = grid_coord((double)i, f.n[d], f.repd[d]) * f.repd[d];
    t -= rint(t);
    double sn, cs;
    sincospi(2.0 * t, &sn, &cs);
    const int off = d == 0 ? 0 : (d == 1 ? f.n[0] : f.n[0] + f.n[1]);
    f.z[off + i] = make_double2(cs, -sn);
}

// B: one work-item per (row, k); sum_{mc = -Mc}^{Mc} A_mc z^mc = z^-Mc Horner(z)
__global__ void grid_fourier_c_kernel(GridFourierArgs f) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n3 = f.n[2], rowlen = 2 * f.Mc + 1;
    const long long n_rows = (long long)(f.Ma + 1) * (2 * f.Mb + 1);
    if (id >= n_rows * n3) return;
    const long long row = id / n3;
    const int k = (int)(id - row * n3);
    const double2 z = f.z[f.n[0] + f.n[1] + k];
    const double2* a = f.A + row * rowlen;
    double2 h = make_double2(0.0, 0.0);
    for (int m = rowlen - 1; m >= 0; --m) h = cfma(h, z, a[m]);
    const double2 zc = make_double2(z.x, -z.y);
    double2 p = make_double2(1.0, 0.0);
    for (int q = 0; q < f.Mc; ++q) p = cmul(p, zc);
    f.B[id] = cmul(h, p);
}

// C: one work-item per (ma, j, k); sum_{mb = -Mb}^{Mb} B_mb z^mb = z^-Mb Horner(z)
__global__ void grid_fourier_b_kernel(GridFourierArgs f) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n2 = f.n[1], n3 = f.n[2], nb = 2 * f.Mb + 1;
    if (id >= (long long)(f.Ma + 1) * n2 * n3) return;
    const int k = (int)(id % n3);
    const long long r = id / n3;
    const int j = (int)(r % n2);
    const int ma = (int)(r / n2);
    const double2 z = f.z[f.n[0] + j];
    double2 h = make_double2(0.0, 0.0);
    for (int m = nb - 1; m >= 0; --m) h = cfma(h, z, f.B[((long long)ma * nb + m) * n3 + k]);
    const double2 zc = make_double2(z.x, -z.y);
    double2 p = make_double2(1.0, 0.0);
    for (int q = 0; q < f.Mb; ++q) p = cmul(p, zc);
    f.C[id] = cmul(h, p);
}

// lr per grid point: Re sum_{ma = 0}^{Ma} C_ma z^ma (Horner), point i fastest
__global__ void grid_fourier_a_kernel(GridFourierArgs f) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = f.n[0], n2 = f.n[1], n3 = f.n[2];
    if (p >= (long long)n1 * n2 * n3) return;
    const int i = (int)(p % n1);
    const long long jk = p / n1;  // j + n2 k
    const double2 z = f.z[i];
    const long long plane = (long long)n2 * n3;
    const long long jk_c = (jk % n2) * n3 + jk / n2;  // j n3 + k
    double2 h = make_double2(0.0, 0.0);
    for (int ma = f.Ma; ma >= 0; --ma) h = cfma(h, z, f.C[ma * plane + jk_c]);
    f.out_lr[f.slot_base + p] = h.x;
}

// -------------------------------------------------------------- combine ----
// Per insertion: E_vdw = sum_s lj_s, E_es = sum_s q_s (sr_s / (4 pi eps0) + lr_s),
// Boltzmann weight and the workgroup's fixed-order partial sums.
__global__ __launch_bounds__(WG) void widom_combine_kernel(GroupArgs g) {
    const SegArgs sg = g.segs[find_segment<SEG_BY_WG>(g.segs, g.n_seg, blockIdx.x)];
    const RunArgs& a = g.runs[sg.run];
    const int lw = (int)blockIdx.x - sg.wg_base, w = sg.w0 + lw;
    const int b = w / a.chunks_per_block, c = w - b * a.chunks_per_block;
    const long long i = (long long)c * WG + threadIdx.x;
    const bool valid = i < a.n_per_block;
    const int ns = a.nsite;
    const long long p0 = sg.slot_base + ((long long)lw * WG + threadIdx.x) * ns;
    // a blocked insertion emitted no point: its per-slot sums were never written
    const bool blocked = valid && a.acc != nullptr && g.pos[p0].w == 0.0;
    double vdw = 0.0, es = 0.0;
    if (blocked) {
        vdw = INFINITY;
        es = INFINITY;
    } else if (valid) {
        for (int s = 0; s < ns; ++s) {
            vdw += g.out_lj[p0 + s];
            if (a.esm != 0) es += a.site_q[s] * (g.out_sr[p0 + s] * a.inv4pe + g.out_lr[p0 + s]);
        }
    }
    const double wgt = exp(-(vdw + es) / a.temperature);
    const bool fin = !isinf(vdw) && !isinf(es);
    if (a.out_e != nullptr && valid) {
        const long long gi = (long long)b * a.n_per_block + i;
        a.out_e[2 * gi] = vdw;
        a.out_e[2 * gi + 1] = es;
    }
    block_sum3(valid ? wgt : 0.0, (valid && fin) ? wgt * vdw : 0.0, (valid && fin) ? wgt * es : 0.0,
               a.partials + 3LL * w);
}

__global__ __launch_bounds__(WG) void grid_combine_kernel(GroupArgs g) {
    const long long p0 = (long long)blockIdx.x * WG;
    const SegArgs sg = g.segs[find_segment<SEG_BY_SLOT>(g.segs, g.n_seg, p0)];
    const RunArgs& a = g.runs[sg.run];
    const long long p = p0 + threadIdx.x, lp = p - sg.slot_base;
    if (lp >= sg.n_slots) return;
    double e = g.out_lj[p];
    if (a.esm != 0) e += a.grid_q * (g.out_sr[p] * a.inv4pe + g.out_lr[p]);
    if (a.grid_f != nullptr) a.grid_f[lp] = (float)e;
    else a.grid_d[lp] = e;
}

// Fixed-order workgroup sum of K values; thread k < K writes the k-th sum to
// out_a[k] (k < 3) or out_b[k - 3].  For k < 3 the arithmetic is block_sum3's.
template <int K>
__device__ inline void block_sum_k(const double* v, double* out_a, double* out_b) {
    __shared__ double red[K][WG / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double sk = wave_sum(v[k]);
        if (lane == 0) red[k][wave] = sk;
    }
    __syncthreads();
    if (tid < K) {
        double t = 0.0;
        for (int w = 0; w < WG / 64; ++w) t += red[tid][w];
        if (tid < 3) out_a[tid] = t;
        else out_b[tid - 3] = t;
    }
}

// Fixed-order per-block reduction of the workgroup partials: one workgroup
// per Widom block of every run, strided partial sums then a shuffle tree.
__global__ __launch_bounds__(WG) void block_sums_kernel(const RunArgs* runs, int n_runs, double* out) {
    int lo = 0, hi = n_runs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (runs[mid].block_base <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const RunArgs& a = runs[lo];
    const int b = (int)blockIdx.x - a.block_base, chunks = a.chunks_per_block;
    if (a.pstride == 3) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int c = threadIdx.x; c < chunks; c += WG) {
            const double* p = a.partials + 3LL * ((long long)b * chunks + c);
            t0 += p[0]; t1 += p[1]; t2 += p[2];
        }
        block_sum3(t0, t1, t2, out + 3LL * blockIdx.x);
    } else {
        double t[GK];
#pragma unroll
        for (int k = 0; k < GK; ++k) t[k] = 0.0;
        for (int c = threadIdx.x; c < chunks; c += WG) {
            const double* p = a.partials + (long long)GK * ((long long)b * chunks + c);
#pragma unroll
            for (int k = 0; k < GK; ++k) t[k] += p[k];
        }
        block_sum_k<GK>(t, out + 3LL * blockIdx.x, a.gsums + (long long)(3 * GT) * b);
    }
}

// ------------------------------------------------------------- adjoint ----
// Per-type Fourier coefficients for the charge derivatives: one wave per kept
// k-vector, lanes over the unit-cell atoms (xyz, w = compact type), unit
// charges.  box_t[entry * GT + t] = wt N_R sum_{atoms of type t} e^{i 2 pi m.x}.
__global__ __launch_bounds__(64) void type_structure_factor_kernel(const double4* ua, long long n, const KVec* kv,
                                                                   int Mb, int Mc, double2* box_t) {
    const int k = blockIdx.x, lane = threadIdx.x;
    const KVec v = kv[k];
    double sre[GT], sim[GT];
#pragma unroll
    for (int t = 0; t < GT; ++t) { sre[t] = 0.0; sim[t] = 0.0; }
    for (long long c = lane; c < n; c += 64) {
        const double4 p = ua[c];
        double ph = v.ma * p.x + v.mb * p.y + v.mc * p.z;
        ph -= rint(ph);
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        const int ty = (int)p.w;
#pragma unroll
        for (int t = 0; t < GT; ++t) {
            sre[t] += (ty == t) ? cs : 0.0;
            sim[t] += (ty == t) ? sn : 0.0;
        }
    }
    const long long idx =
        ((long long)v.ma * (2 * Mb + 1) + (long long)v.mb + Mb) * (2 * Mc + 1) + (long long)v.mc + Mc;
#pragma unroll
    for (int t = 0; t < GT; ++t) {
        const double r = wave_sum(sre[t]), i = wave_sum(sim[t]);
        if (lane == 0) box_t[idx * GT + t] = make_double2(v.wnr * r, v.wnr * i);
    }
}

// Per-type Fourier sums of every charged point of a gradient run, FP64:
// F_t(u) = sum over the kept m of Re(A_t(m) e^{-i 2 pi m.u}) for NT framework
// types.  Each (ma, mb) row runs over its kept mc range [lo, hi] directly: the
// phase z_a^ma z_b^mb z_c^mc steps along the row by z_c, and z_c^lo is carried
// from row to row by single steps of z_c or its conjugate (the kept ranges of
// neighbouring rows differ by a few entries).  Per kept entry and point that
// is 2 NT FMAs for the sums and one complex product for the phase; the
// coefficients are read by scalar loads shared by the FOURIER_PTS points of a
// work-item, as in point_fourier_kernel.
template <int NT>
__global__ __launch_bounds__(FOURIER_WG) void point_fourier_grad_kernel(GroupArgs g) {
    const int tid = threadIdx.x;
    const long long p0 = (long long)blockIdx.x * WG;  // WG slots per workgroup, FOURIER_PTS per work-item
    const SegArgs sg = g.segs[find_segment<SEG_BY_SLOT>(g.segs, g.n_seg, p0)];
    const RunArgs& a = g.runs[sg.run];
    if (a.esm == 0) return;  // uniform: the whole workgroup leaves
    FourierPoint P[FOURIER_PTS];
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k) P[k] = fourier_point(g, sg, a, p0 + tid + k * FOURIER_WG);
    // wave-uniform: scalar loop control (the loads are not provably uniform)
    const int Ma = __builtin_amdgcn_readfirstlane(a.Ma), Mb = __builtin_amdgcn_readfirstlane(a.Mb),
              Mc = __builtin_amdgcn_readfirstlane(a.Mc);
    double2 Eb_neg[FOURIER_PTS], Pa[FOURIER_PTS], Pb[FOURIER_PTS], Zc[FOURIER_PTS], Z[FOURIER_PTS];
    double acc[FOURIER_PTS][NT];
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k) {
        Eb_neg[k] = make_double2(1.0, 0.0);
        const double2 Eb_c = make_double2(P[k].Eb.x, -P[k].Eb.y);
        for (int q = 0; q < Mb; ++q) Eb_neg[k] = cmul(Eb_neg[k], Eb_c);
        Pa[k] = make_double2(1.0, 0.0);
        Pb[k] = Eb_neg[k];
        Zc[k] = make_double2(P[k].zr, P[k].zi);
        Z[k] = make_double2(1.0, 0.0);  // z_c^zc_at
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[k][t] = 0.0;
    }
    int zc_at = 0;
    const int rowlen = 2 * Mc + 1, nb = 2 * Mb + 1;
    const int n_rows = (Ma + 1) * nb;
    const gconst_ptr<double> At = uniform_ptr(reinterpret_cast<const double*>(a.A_t));
    const gconst_ptr<int> krows = uniform_ptr(reinterpret_cast<const int*>(a.krows));
    for (int r = 0; r < n_rows; ++r) {
        if (r > 0 && (r % nb) == 0) {  // next m_a
#pragma unroll
            for (int k = 0; k < FOURIER_PTS; ++k) {
                Pa[k] = cmul(Pa[k], P[k].Ea);
                Pb[k] = cmul(Pa[k], Eb_neg[k]);
            }
        }
        const int lo = krows[2 * r], hi = krows[2 * r + 1];
        if (lo <= hi) {
            for (; zc_at < lo; ++zc_at)
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) Z[k] = cmul(Z[k], Zc[k]);
            for (; zc_at > lo; --zc_at)
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) Z[k] = cmul(Z[k], make_double2(Zc[k].x, -Zc[k].y));
            double2 Q[FOURIER_PTS];
#pragma unroll
            for (int k = 0; k < FOURIER_PTS; ++k) Q[k] = cmul(Pb[k], Z[k]);
            // entry (r, mc): 2 GT doubles from At + 2 GT ((r rowlen) + mc + Mc)
            const gconst_ptr<double> row = At + 2LL * GT * ((long long)r * rowlen + Mc);
#pragma unroll 2
            for (int mc = lo; mc <= hi; ++mc) {
                const gconst_ptr<double> e = row + 2LL * GT * mc;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const double Ar = e[2 * t], Ai = e[2 * t + 1];
#pragma unroll
                    for (int k = 0; k < FOURIER_PTS; ++k) acc[k][t] = fma(Ar, Q[k].x, fma(-Ai, Q[k].y, acc[k][t]));
                }
#pragma unroll
                for (int k = 0; k < FOURIER_PTS; ++k) Q[k] = cmul(Q[k], Zc[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < FOURIER_PTS; ++k) Pb[k] = cmul(Pb[k], P[k].Eb);
    }
#pragma unroll
    for (int k = 0; k < FOURIER_PTS; ++k) {
        if (P[k].valid) {
            double* o = g.out_g + 3 * GT * g.g_stride + P[k].pid;
#pragma unroll
            for (int t = 0; t < NT; ++t) o[t * g.g_stride] = P[k].charged ? acc[k][t] : 0.0;
        }
    }
}

// Per insertion of a gradient run: the energies and weight of
// widom_combine_kernel, and dE/d(eps_t, sigma_t, q_t) from the per-type sums
// through the mixing-rule Jacobian:
//   dE/d eps_t   = sum_s 4 RE_st d eps_ts/d eps_t   + RS_st d sigma_ts/d eps_t
//   dE/d sigma_t = sum_s 4 RE_st d eps_ts/d sigma_t + RS_st d sigma_ts/d sigma_t
//   dE/d q_t     = sum_s q_s (RQ_st / (4 pi eps0) + LR_st)
// Workgroup partials: (S, W_vdw, W_es, then w dE/d theta for every type).
// Two passes over the workgroup's WG nsite contiguous slots.  First a
// work-item per insertion: energies, weight and the (S, W) terms, exactly as
// widom_combine_kernel, with the weight (or -1 for an insertion that adds
// nothing) left in LDS.  Then a work-item per slot, lane-contiguous so the
// 4 NT per-type sum arrays are read coalesced: slot j = insertion j / nsite,
// site j % nsite adds w (Jacobian . sums) for every type.  NT >= every run's
// type count (the group's largest); types a run lacks add nothing.
template <int NT>
__global__ __launch_bounds__(WG) void widom_combine_grad_kernel(GroupArgs g) {
    __shared__ double s_w[WG];
    __shared__ double4 s_jac[NT * PM_MAX_SITES];
    __shared__ double s_q[PM_MAX_SITES];
    const int tid = threadIdx.x;
    const SegArgs sg = g.segs[find_segment<SEG_BY_WG>(g.segs, g.n_seg, blockIdx.x)];
    const RunArgs& a = g.runs[sg.run];
    const int lw = (int)blockIdx.x - sg.wg_base, w = sg.w0 + lw;
    const int b = w / a.chunks_per_block, c = w - b * a.chunks_per_block;
    const long long i = (long long)c * WG + tid;
    const bool valid = i < a.n_per_block;
    const int ns = a.nsite, n_gt = a.n_gt;
    const long long q0 = sg.slot_base + (long long)lw * WG * ns;  // the workgroup's first slot
    const long long p0 = q0 + (long long)tid * ns;
    if (tid < n_gt * ns) s_jac[tid] = a.jac[tid];
    if (tid < ns) s_q[tid] = a.site_q[tid];
    double vdw = 0.0, es = 0.0;
    const bool blocked = valid && a.acc != nullptr && g.pos[p0].w == 0.0;
    if (blocked) {
        vdw = INFINITY;
        es = INFINITY;
    } else if (valid) {
#pragma unroll
        for (int s = 0; s < PM_MAX_SITES; ++s) {
            if (s >= ns) break;
            vdw += g.out_lj[p0 + s];
            if (a.esm != 0) es += a.site_q[s] * (g.out_sr[p0 + s] * a.inv4pe + g.out_lr[p0 + s]);
        }
    }
    const double wgt = exp(-(vdw + es) / a.temperature);
    const bool fin = !isinf(vdw) && !isinf(es);
    const bool use = valid && fin;
    if (a.out_e != nullptr && valid) {
        const long long gi = (long long)b * a.n_per_block + i;
        a.out_e[2 * gi] = vdw;
        a.out_e[2 * gi + 1] = es;
    }
    s_w[tid] = use ? wgt : -1.0;
    double v[GK];
    v[0] = valid ? wgt : 0.0;
    v[1] = use ? wgt * vdw : 0.0;
    v[2] = use ? wgt * es : 0.0;
    double de[GT], dsg[GT], dq[GT];
#pragma unroll
    for (int t = 0; t < GT; ++t) { de[t] = 0.0; dsg[t] = 0.0; dq[t] = 0.0; }
    __syncthreads();
    // (every sum array exists for GT types; those a run lacks hold zeros or,
    // for a run without charges, are not used)
    const long long st = g.g_stride;
    const bool charged = a.esm != 0;
    for (int j = tid; j < WG * ns; j += WG) {
        const int ins = j / ns, s = j - ins * ns;
        const double wj = s_w[ins];
        if (wj < 0.0) continue;
        const double* gs = g.out_g + (q0 + j);
        double w0[NT], w1[NT], w2[NT], w3[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            w0[t] = gs[t * st];
            w1[t] = gs[(GT + t) * st];
            w2[t] = charged ? gs[(2 * GT + t) * st] : 0.0;
            w3[t] = charged ? gs[(3 * GT + t) * st] : 0.0;
        }
        const double qs = s_q[s];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t < n_gt) {
                const double4 J = s_jac[t * ns + s];
                const double re = 4.0 * w0[t], rs = w1[t];
                de[t] += wj * (re * J.x + rs * J.z);
                dsg[t] += wj * (re * J.y + rs * J.w);
                if (charged) dq[t] += wj * (qs * (w2[t] * a.inv4pe + w3[t]));
            }
        }
    }
#pragma unroll
    for (int t = 0; t < GT; ++t) {
        v[3 + 3 * t] = de[t];
        v[4 + 3 * t] = dsg[t];
        v[5 + 3 * t] = dq[t];
    }
    double* part = a.partials + (long long)GK * w;
    block_sum_k<GK>(v, part, part + 3);
}

// ----------------------------------------------------------- measurement ----
// Plain device copy (16-byte vector loads and stores, grid-stride): the
// calibration launch for HBM traffic counters, n16 = bytes / 16.
__global__ __launch_bounds__(WG) void copy_probe_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                         long long n16) {
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n16; i += (long long)gridDim.x * WG)
        dst[i] = src[i];
}

}  // namespace pmk
