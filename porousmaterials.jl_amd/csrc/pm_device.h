// pm_device.h -- gfx950 (CDNA4, wave64) device code of libpmhip.so.
//
// Numerical contract (restating the reference, PorousMaterials.jl v0.4.2):
//   LJ pair        src/vdw_energetics.jl:19-22    4 eps s^3 (s^3 - 1), s = sigma^2/r^2
//   pair loop      src/vdw_energetics.jl:25-47    nearest image, strict r^2 < rc^2,
//                                                 overlap r^2 < 0.1 -> +Inf
//   real space     src/electrostatics.jl:377-386  q erfc(alpha r) / (4 pi eps0 r), r < rc
//   Fourier        src/electrostatics.jl:354-372  sum_k wt_k sum_c q_c cos(k.(x_c - x)),
//                                                 here from the unit-cell structure
//                                                 factor (the same sum, reorganised)
//   Widom          src/henry.jl:212-243           per-block sum e^{-E/T}, E e^{-E/T}
//   energy grid    src/grid.jl:313-355            point (i,j,k)/(n-1) ./ rep
//
// Two execution schemes share the pair arithmetic:
//   pose_kernel       one work-item per pose (all of its sites), every framework
//                     atom streamed through LDS, nearest image per pair; serves the
//                     explicit-pose API (vdw_energy, electrostatic_potential_energy).
//   point engine      every probe site is an independent point.  Points are grouped
//                     into spatially compact chunks (one cell of a counting sort, or
//                     one brick of a grid) of <= 256 per workgroup.  A chunk scans an
//                     image list of the framework that covers every image within
//                     r_c + chunk radius of any cell, keeps those images (Cartesian,
//                     FP64, relative to the chunk centre) in LDS and evaluates only
//                     them: no per-pair wrap or basis change, ~1/4 of the pairs.
//                     The Fourier part runs in its own FP64 kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmhip.h"

namespace pmk {

constexpr int WG = 256;      // work-items per workgroup (4 waves)
constexpr int TILE = 256;    // framework atoms per LDS tile
constexpr int KTILE = 128;   // k-vectors per LDS tile (pose kernel)
constexpr int FT = 2048;     // Fourier coefficients per LDS tile (point engine)
constexpr double R2_OVERLAP = 0.1;  // src/vdw_energetics.jl:2

enum { MODE_WIDOM = 0, MODE_GRID = 2 };

struct KVec {
    double ma, mb, mc, are, aim, pad;
};

struct Site {
    double ox, oy, oz;  // Cartesian offset from the centre of mass (template)
    double q;           // point charge (e), 0 if none
    int lj;             // 1 if an LJ sphere
    int pad;
};

template <typename Real>
struct LJP {
    Real sig2, eps4;
};

// ---------------------------------------------------------------- RNG ----
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Insertion g draws three
// blocks: counter (g_lo, g_hi, w, 0x57494430), key (seed_lo, seed_hi), w = 0,1,2.
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ inline void insertion_uniforms(unsigned long long seed, unsigned long long g, double u[6]) {
#pragma unroll
    for (uint32_t w = 0; w < 3; ++w) {
        uint32_t c[4] = {(uint32_t)g, (uint32_t)(g >> 32), w, 0x57494430u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        u[2 * w] = u53(c[0], c[1]);
        u[2 * w + 1] = u53(c[2], c[3]);
    }
}

// Arvo rotation (src/molecule.jl:233-244): R = -H Rz(2 pi u0), H = I - 2 v v^T,
// v = (cos(2 pi u1) sqrt(u2), sin(2 pi u1) sqrt(u2), sqrt(1 - u2)); column-major.
__device__ inline void arvo_rotation(const double* u, double R[9]) {
    double s1, c1, s2, c2;
    sincospi(2.0 * u[0], &s1, &c1);
    sincospi(2.0 * u[1], &s2, &c2);
    double sq = sqrt(u[2]);
    double v0 = c2 * sq, v1 = s2 * sq, v2 = sqrt(1.0 - u[2]);
    double H[9] = {1.0 - 2.0 * v0 * v0, -2.0 * v1 * v0,       -2.0 * v2 * v0,
                   -2.0 * v0 * v1,       1.0 - 2.0 * v1 * v1, -2.0 * v2 * v1,
                   -2.0 * v0 * v2,       -2.0 * v1 * v2,       1.0 - 2.0 * v2 * v2};
    // Rz columns: (c1, -s1, 0), (s1, c1, 0), (0, 0, 1)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        R[r] = -(H[r] * c1 - H[r + 3] * s1);
        R[r + 3] = -(H[r] * s1 + H[r + 3] * c1);
        R[r + 6] = -H[r + 6];
    }
}

// ---------------------------------------------------------- pair maths ----
// FP32 erfc: Numerical Recipes "erfcc" Chebyshev fit (Press et al., Numerical
// Recipes in C, 2nd ed., section 6.2), fractional error < 1.2e-7 for x >= 0.
__device__ inline float erfc_f32(float x) {
    float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
    float p = 0.17087277f;
    p = fmaf(p, t, -0.82215223f);
    p = fmaf(p, t, 1.48851587f);
    p = fmaf(p, t, -1.13520398f);
    p = fmaf(p, t, 0.27886807f);
    p = fmaf(p, t, -0.18628806f);
    p = fmaf(p, t, 0.09678418f);
    p = fmaf(p, t, 0.37409196f);
    p = fmaf(p, t, 1.00002368f);
    p = fmaf(p, t, -1.26551223f);
    float y = p - x * x;  // exponent, natural-log units
    return t * __builtin_amdgcn_exp2f(y * 1.4426950408889634f);
}

template <typename Real>
struct PairMath;

template <>
struct PairMath<float> {
    // 4 eps s^3 (s^3 - 1), s = sigma^2 / r^2
    __device__ static inline float lj(double r2, float sig2, float eps4) {
        float s = sig2 * __builtin_amdgcn_rcpf((float)r2);
        float s3 = s * s * s;
        return eps4 * fmaf(s3, s3, -s3);
    }
    // q erfc(alpha r) / r
    __device__ static inline float coul(double r2, double q, double alpha) {
        float rf2 = (float)r2;
        float rinv = __builtin_amdgcn_rsqf(rf2);
        float r = rf2 * rinv;
        return (float)q * erfc_f32((float)alpha * r) * rinv;
    }
};

template <>
struct PairMath<double> {
    __device__ static inline double lj(double r2, double sig2, double eps4) {
        double s = sig2 / r2;
        double s3 = s * s * s;
        return eps4 * s3 * (s3 - 1.0);
    }
    __device__ static inline double coul(double r2, double q, double alpha) {
        double r = sqrt(r2);
        return q / r * erfc(r * alpha);
    }
};

// --------------------------------------------------------- reductions ----
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed-order workgroup sum of three values into out[0..2] (thread 0 writes).
__device__ inline void block_sum3(double v0, double v1, double v2, double* out) {
    __shared__ double red[3][WG / 64];
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (lane == 0) { red[0][wave] = v0; red[1][wave] = v1; red[2][wave] = v2; }
    __syncthreads();
    if (tid == 0) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int w = 0; w < WG / 64; ++w) { t0 += red[0][w]; t1 += red[1][w]; t2 += red[2][w]; }
        out[0] = t0; out[1] = t1; out[2] = t2;
    }
}

// =========================================================================
// Explicit poses: one work-item per pose, NS sites each.
// =========================================================================
struct PoseArgs {
    int ns, n_atoms, n_types, nk;
    long long n_items;
    double f11, f12, f13, f22, f23, f33;  // replicated-box f_to_c (upper triangular)
    double repd[3];                       // replication factors (unit-cell phases)
    double r2_lj, r2_sr, alpha, inv4pe;
    Site site[PM_MAX_SITES];
    const double4* atoms;  // replicated atoms: fractional xyz + charge
    const int* species;
    const void* ljtab;     // [n_types][ns] of LJP<Real>
    const KVec* kv;
    const double* pose_xf; // n_items x ns x 3
    double* out_e;         // n_items x 3: (vdw, sr, lr)
};

template <int NS, bool ES, typename Real>
__global__ __launch_bounds__(WG) void pose_kernel(PoseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int ATOM_BYTES = TILE * (int)sizeof(double4);
    constexpr int K_BYTES = KTILE * (int)sizeof(KVec);
    constexpr int REGION = ATOM_BYTES > K_BYTES ? ATOM_BYTES : K_BYTES;
    double4* s_atoms = reinterpret_cast<double4*>(smem);
    KVec* s_k = reinterpret_cast<KVec*>(smem);
    int* s_type = reinterpret_cast<int*>(smem + REGION);
    LJP<Real>* s_lj = reinterpret_cast<LJP<Real>*>(smem + REGION + TILE * sizeof(int));

    const int tid = threadIdx.x;
    const LJP<Real>* gtab = reinterpret_cast<const LJP<Real>*>(a.ljtab);
    for (int i = tid; i < a.n_types * NS; i += WG) s_lj[i] = gtab[i];

    const long long item = (long long)blockIdx.x * WG + tid;
    const bool valid = item < a.n_items;
    const long long it = valid ? item : 0;
    double sx[NS], sy[NS], sz[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double* p = a.pose_xf + (it * NS + s) * 3;
        sx[s] = p[0]; sy[s] = p[1]; sz[s] = p[2];
    }

    double acc_lj[NS], acc_sr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { acc_lj[s] = 0.0; acc_sr[s] = 0.0; }
    bool ovl = false;
    const double f11 = a.f11, f12 = a.f12, f13 = a.f13, f22 = a.f22, f23 = a.f23, f33 = a.f33;
    const double r2lj = a.r2_lj, r2sr = a.r2_sr, alpha = a.alpha;
    for (int base = 0; base < a.n_atoms; base += TILE) {
        int nt = min(TILE, a.n_atoms - base);
        __syncthreads();
        if (tid < nt) {
            s_atoms[tid] = a.atoms[base + tid];
            s_type[tid] = a.species[base + tid];
        }
        __syncthreads();
        for (int j = 0; j < nt; ++j) {
            const double4 at = s_atoms[j];
            const int t = s_type[j];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double dx = at.x - sx[s], dy = at.y - sy[s], dz = at.z - sz[s];
                dx -= rint(dx);
                dy -= rint(dy);
                dz -= rint(dz);
                double cx = f11 * dx + f12 * dy + f13 * dz;
                double cy = f22 * dy + f23 * dz;
                double cz = f33 * dz;
                double r2 = cx * cx + cy * cy + cz * cz;
                const LJP<Real> p = s_lj[t * NS + s];
                bool in = r2 < r2lj;
                ovl |= in & (r2 < R2_OVERLAP) & (a.site[s].lj != 0);
                Real e = PairMath<Real>::lj(r2, p.sig2, p.eps4);
                acc_lj[s] += in ? (double)e : 0.0;
                if (ES) {
                    bool ins = (r2 < r2sr) & (at.w != 0.0);
                    Real c = PairMath<Real>::coul(r2, at.w, alpha);
                    acc_sr[s] += ins ? (double)c : 0.0;
                }
            }
        }
    }

    double acc_lr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc_lr[s] = 0.0;
    if (ES) {
        double ux[NS], uy[NS], uz[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            ux[s] = sx[s] * a.repd[0];
            uy[s] = sy[s] * a.repd[1];
            uz[s] = sz[s] * a.repd[2];
        }
        for (int base = 0; base < a.nk; base += KTILE) {
            int nt = min(KTILE, a.nk - base);
            __syncthreads();
            if (tid < nt) s_k[tid] = a.kv[base + tid];
            __syncthreads();
            for (int k = 0; k < nt; ++k) {
                const KVec kv = s_k[k];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (a.site[s].q != 0.0) {
                        double t = kv.ma * ux[s] + kv.mb * uy[s] + kv.mc * uz[s];
                        t -= rint(t);
                        double sn, cs;
                        sincospi(2.0 * t, &sn, &cs);
                        acc_lr[s] += kv.are * cs + kv.aim * sn;
                    }
                }
            }
        }
    }

    double vdw = 0.0, sr = 0.0, lr = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        vdw += acc_lj[s];
        sr += a.site[s].q * acc_sr[s];
        lr += a.site[s].q * acc_lr[s];
    }
    if (ovl) vdw = INFINITY;
    if (valid) {
        a.out_e[3 * item + 0] = vdw;
        a.out_e[3 * item + 1] = sr * a.inv4pe;
        a.out_e[3 * item + 2] = lr;
    }
}

// Unit-cell structure factor per kept k-vector: one wave per k-vector,
// lanes stride over the unit-cell charges, FP64 phases and sums.
__global__ __launch_bounds__(64) void structure_factor_kernel(const double4* uq, long long n,
                                                              const int* m, const double* wt,
                                                              double nr, KVec* out) {
    int k = blockIdx.x, lane = threadIdx.x;
    double ma = m[3 * k], mb = m[3 * k + 1], mc = m[3 * k + 2];
    double sre = 0.0, sim = 0.0;
    for (long long c = lane; c < n; c += 64) {
        double4 p = uq[c];
        double t = ma * p.x + mb * p.y + mc * p.z;
        t -= rint(t);
        double sn, cs;
        sincospi(2.0 * t, &sn, &cs);
        sre += p.w * cs;
        sim += p.w * sn;
    }
    sre = wave_sum(sre);
    sim = wave_sum(sim);
    if (lane == 0) {
        KVec kv;
        kv.ma = ma; kv.mb = mb; kv.mc = mc;
        kv.are = wt[k] * nr * sre;
        kv.aim = wt[k] * nr * sim;
        kv.pad = 0.0;
        out[k] = kv;
    }
}

// Dense Fourier box from the kept k-vectors: A[(ma*(2Mb+1) + mb+Mb)*(2Mc+1) + mc+Mc].
__global__ void fourier_box_kernel(const KVec* kv, int nk, int Mb, int Mc, double2* box) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    KVec v = kv[k];
    long long idx = ((long long)v.ma * (2 * Mb + 1) + (long long)v.mb + Mb) * (2 * Mc + 1) + (long long)v.mc + Mc;
    box[idx] = make_double2(v.are, v.aim);
}

// =========================================================================
// Point engine
// =========================================================================
struct GenArgs {
    unsigned long long seed;
    int rotate, nsite;
    int n_blocks, chunks_per_block;
    long long n_per_block;
    int w0;                       // first insertion workgroup of this batch
    int ncell[3];
    double off[PM_MAX_SITES][3];  // site offsets (Cartesian, from the COM)
    double c2f[9];                // replicated-box c_to_f, column-major
    double4* pos;                 // per point slot: fractional xyz (wrapped), w = 1 valid
    unsigned* cell;               // per point slot: cell id or 0xffffffff
    unsigned* count;              // per cell
};

// One work-item per insertion slot: pose from the stream, sites wrapped into
// [0,1)^3 (integer shifts change neither the nearest-image pairs nor the
// Fourier phases m.(R o x)), binned into cells.
__global__ __launch_bounds__(WG) void widom_gen_kernel(GenArgs a) {
    const int w = a.w0 + blockIdx.x;
    const int b = w / a.chunks_per_block, c = w - b * a.chunks_per_block;
    const long long i = (long long)c * WG + threadIdx.x;
    const bool valid = i < a.n_per_block;
    const long long slot = (long long)blockIdx.x * WG + threadIdx.x;
    if (!valid) {
        for (int s = 0; s < a.nsite; ++s) {
            a.pos[slot * a.nsite + s] = make_double4(0.0, 0.0, 0.0, 0.0);
            a.cell[slot * a.nsite + s] = 0xffffffffu;
        }
        return;
    }
    const unsigned long long g = (unsigned long long)b * a.n_per_block + i;
    double u[6];
    insertion_uniforms(a.seed, g, u);
    double R[9];
    if (a.rotate) arvo_rotation(&u[3], R);
    for (int s = 0; s < a.nsite; ++s) {
        double y0 = a.off[s][0], y1 = a.off[s][1], y2 = a.off[s][2];
        if (a.rotate) {
            double t0 = R[0] * y0 + R[3] * y1 + R[6] * y2;
            double t1 = R[1] * y0 + R[4] * y1 + R[7] * y2;
            double t2 = R[2] * y0 + R[5] * y1 + R[8] * y2;
            y0 = t0; y1 = t1; y2 = t2;
        }
        double x = u[0] + (a.c2f[0] * y0 + a.c2f[3] * y1 + a.c2f[6] * y2);
        double y = u[1] + (a.c2f[1] * y0 + a.c2f[4] * y1 + a.c2f[7] * y2);
        double z = u[2] + (a.c2f[2] * y0 + a.c2f[5] * y1 + a.c2f[8] * y2);
        x -= floor(x);
        y -= floor(y);
        z -= floor(z);
        int ia = min((int)(x * a.ncell[0]), a.ncell[0] - 1);
        int ib = min((int)(y * a.ncell[1]), a.ncell[1] - 1);
        int ic = min((int)(z * a.ncell[2]), a.ncell[2] - 1);
        unsigned cid = (unsigned)(ia + a.ncell[0] * (ib + a.ncell[1] * ic));
        a.pos[slot * a.nsite + s] = make_double4(x, y, z, 1.0);
        a.cell[slot * a.nsite + s] = cid;
        atomicAdd(&a.count[cid], 1u);
    }
}

__global__ void fill_u32_kernel(unsigned* p, long long n, unsigned v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive scans over cells of the point counts and of the slice counts
// (ceil(count / WG)); one workgroup of 1024 work-items walks the cells.
__global__ __launch_bounds__(1024) void cell_scan_kernel(const unsigned* count, unsigned* point_off,
                                                         unsigned* slice_off, unsigned* n_slices,
                                                         int n_cells) {
    __shared__ unsigned s_p[1024 / 64], s_s[1024 / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned run_p = 0, run_s = 0;
    for (int base = 0; base < n_cells; base += 1024) {
        int c = base + tid;
        unsigned v = c < n_cells ? count[c] : 0u;
        unsigned sv = (v + WG - 1) / WG;
        // inclusive wave scans
        unsigned ip = v, is = sv;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            unsigned tp = __shfl_up(ip, off, 64), ts = __shfl_up(is, off, 64);
            if (lane >= off) { ip += tp; is += ts; }
        }
        if (lane == 63) { s_p[wave] = ip; s_s[wave] = is; }
        __syncthreads();
        unsigned wp = 0, ws = 0, tp = 0, ts = 0;
        for (int w = 0; w < 1024 / 64; ++w) {
            if (w < wave) { wp += s_p[w]; ws += s_s[w]; }
            tp += s_p[w];
            ts += s_s[w];
        }
        if (c < n_cells) {
            point_off[c] = run_p + wp + ip - v;
            slice_off[c] = run_s + ws + is - sv;
        }
        run_p += tp;
        run_s += ts;
        __syncthreads();
    }
    if (tid == 0) {
        point_off[n_cells] = run_p;
        slice_off[n_cells] = run_s;
        *n_slices = run_s;
    }
}

__global__ void scatter_points_kernel(const unsigned* cell, long long n_pts, const unsigned* point_off,
                                      unsigned* cursor, unsigned* order) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts) return;
    unsigned c = cell[p];
    if (c == 0xffffffffu) return;
    unsigned slot = point_off[c] + atomicAdd(&cursor[c], 1u);
    order[slot] = (unsigned)p;
}

__global__ void slice_items_kernel(const unsigned* count, const unsigned* slice_off, uint2* items,
                                   int n_cells) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    unsigned ns = (count[c] + WG - 1) / WG, base = slice_off[c];
    for (unsigned k = 0; k < ns; ++k) items[base + k] = make_uint2((unsigned)c, k);
}

constexpr int ERFC_SEG = 512;  // cubic segments of erfc(alpha r) on [0, r_sr] (FP32 path)

struct PArgs {
    double f11, f12, f13, f22, f23, f33;  // replicated-box f_to_c (upper triangular)
    double repd[3];
    double r2_lj, r2_sr, alpha;
    double r_cmax;       // max(r_lj, r_sr)
    double chunk_r;      // bound on |point - chunk centre| (Cartesian)
    double erfc_inv_h;   // ERFC_SEG / r_sr
    int fast;           // image list covers r_cmax + chunk_r around every cell
    int mode;            // MODE_WIDOM (cells) or MODE_GRID (bricks)
    int nsite, n_types, n_img, n_atoms;
    unsigned site_lj_mask;
    // cells
    int ncell[3];
    const unsigned* point_off;
    const uint2* items;
    const unsigned* n_items;
    const unsigned* order;
    const double4* pos;
    // bricks
    int grid_n[3], brick[3], nbrick[3];
    // framework
    const double4* img;   // image list: replicated-box fractional xyz (may lie outside [0,1)) + charge
    const int* img_type;
    const double4* atoms; // replicated atoms (fallback path)
    const int* species;
    const void* ljtab;    // [n_types][nsite] LJP<Real>
    const float4* erfc_tab;  // ERFC_SEG cubics of erfc(alpha r) in t = r/h - i (FP32 path)
    double* out_lj;       // by point slot / grid index
    double* out_sr;
    unsigned long long* cand_pairs;  // evaluated candidate pairs (reporting)
};

__device__ inline double grid_coord(double idx, int n, double rep) {
    return (n > 1 ? idx / (double)(n - 1) : 0.0) / rep;
}

// q erfc(alpha r) / r with erfc(alpha r) from the cubic table: segment
// i = floor(r/h), clamped to the table (r beyond r_sr is masked by the
// caller), t = r/h - i in [0, 1).
__device__ inline float coul_table(double r2, double q, float inv_h, const float4* tab) {
    const float rf2 = (float)r2;
    const float rinv = __builtin_amdgcn_rsqf(rf2);
    const float x = fminf(rf2 * rinv * inv_h, (float)(ERFC_SEG - 1));
    const int i = (int)x;
    const float t = x - (float)i;
    const float4 c = tab[i];
    const float e = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    return (float)q * e * rinv;
}

// One candidate pair: LJ into plj, real-space Ewald into psr.
// ESM: 0 no electrostatics, 1 LJ and Ewald share the cutoff, 2 separate cutoffs.
template <int ESM, typename Real>
__device__ inline void pair_term(double r2, double q, const LJP<Real>& p, bool is_lj, double r2lj,
                                 double r2sr, double alpha, float inv_h, const float4* tab, bool& ovl,
                                 Real& plj, Real& psr) {
    const bool in = r2 < r2lj;
    ovl |= in & (r2 < R2_OVERLAP) & is_lj;
    const Real e = PairMath<Real>::lj(r2, p.sig2, p.eps4);
    plj += in ? e : (Real)0;
    if constexpr (ESM != 0) {
        const bool ins = ((ESM == 1) ? in : (r2 < r2sr)) & (q != 0.0);
        Real cv;
        if constexpr (sizeof(Real) == 4) cv = coul_table(r2, q, inv_h, tab);
        else cv = PairMath<Real>::coul(r2, q, alpha);
        psr += ins ? cv : (Real)0;
    }
}

template <int ESM, typename Real>
__global__ __launch_bounds__(WG) void point_pair_kernel(PArgs a) {
    constexpr bool ES = ESM != 0;
    constexpr int ERFC_BYTES = (ES && sizeof(Real) == 4) ? ERFC_SEG * (int)sizeof(float4) : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    double4* s_c = reinterpret_cast<double4*>(smem);
    int* s_t = reinterpret_cast<int*>(smem + TILE * sizeof(double4));
    float4* s_erfc = reinterpret_cast<float4*>(smem + TILE * (sizeof(double4) + sizeof(int)));
    LJP<Real>* s_lj = reinterpret_cast<LJP<Real>*>(smem + TILE * (sizeof(double4) + sizeof(int)) + ERFC_BYTES);
    __shared__ int s_wcnt[WG / 64];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double cf0, cf1, cf2;  // chunk centre (fractional)
    bool valid;
    long long pid;
    double px, py, pz;
    int site;
    if (a.mode == MODE_WIDOM) {
        if (blockIdx.x >= *a.n_items) return;  // the whole workgroup leaves
        const uint2 it = a.items[blockIdx.x];
        const int c = (int)it.x;
        const int ia = c % a.ncell[0], ib = (c / a.ncell[0]) % a.ncell[1], ic = c / (a.ncell[0] * a.ncell[1]);
        cf0 = (ia + 0.5) / a.ncell[0];
        cf1 = (ib + 0.5) / a.ncell[1];
        cf2 = (ic + 0.5) / a.ncell[2];
        const unsigned k = a.point_off[c] + it.y * WG + tid;
        valid = k < a.point_off[c + 1];
        pid = valid ? (long long)a.order[k] : 0;
        const double4 p = valid ? a.pos[pid] : make_double4(cf0, cf1, cf2, 0.0);
        px = p.x; py = p.y; pz = p.z;
        site = (int)(pid % a.nsite);
    } else {
        const int nb0 = a.nbrick[0], nb1 = a.nbrick[1];
        const int bi = blockIdx.x % nb0, bj = (blockIdx.x / nb0) % nb1, bk = blockIdx.x / (nb0 * nb1);
        const int i0 = bi * a.brick[0], j0 = bj * a.brick[1], k0 = bk * a.brick[2];
        const int ti = tid % a.brick[0], tj = (tid / a.brick[0]) % a.brick[1], tk = tid / (a.brick[0] * a.brick[1]);
        const int i = i0 + ti, j = j0 + tj, k = k0 + tk;
        const int n1 = a.grid_n[0], n2 = a.grid_n[1], n3 = a.grid_n[2];
        valid = i < n1 && j < n2 && k < n3;
        pid = valid ? (long long)i + (long long)n1 * ((long long)j + (long long)n2 * k) : 0;
        px = grid_coord(i, n1, a.repd[0]);
        py = grid_coord(j, n2, a.repd[1]);
        pz = grid_coord(k, n3, a.repd[2]);
        cf0 = grid_coord(i0 + 0.5 * (a.brick[0] - 1), n1, a.repd[0]);
        cf1 = grid_coord(j0 + 0.5 * (a.brick[1] - 1), n2, a.repd[1]);
        cf2 = grid_coord(k0 + 0.5 * (a.brick[2] - 1), n3, a.repd[2]);
        site = 0;
    }
    const LJP<Real>* gtab = reinterpret_cast<const LJP<Real>*>(a.ljtab);
    for (int t = tid; t < a.n_types * a.nsite; t += WG) s_lj[t] = gtab[t];
    if constexpr (ERFC_BYTES != 0)
        for (int t = tid; t < ERFC_SEG; t += WG) s_erfc[t] = a.erfc_tab[t];

    const double f11 = a.f11, f12 = a.f12, f13 = a.f13, f22 = a.f22, f23 = a.f23, f33 = a.f33;
    const double r2lj = a.r2_lj, r2sr = a.r2_sr, alpha = a.alpha;
    const float inv_h = (float)a.erfc_inv_h;
    const bool is_lj = (a.site_lj_mask >> site) & 1u;
    const int nsite = a.nsite;
    // a wave with no valid point skips the pair loops; it still takes part in
    // the tile scans and barriers
    const bool wave_live = __ballot(valid) != 0ull;
    double acc_lj = 0.0, acc_sr = 0.0;
    bool ovl = false;
    unsigned long long n_cand_total = 0;

    if (a.fast) {
        // the point relative to the chunk centre (Cartesian, FP64 wrap)
        double dx = px - cf0, dy = py - cf1, dz = pz - cf2;
        dx -= rint(dx);
        dy -= rint(dy);
        dz -= rint(dz);
        const double prx = f11 * dx + f12 * dy + f13 * dz;
        const double pry = f22 * dy + f23 * dz;
        const double prz = f33 * dz;
        const double Rc = a.r_cmax + a.chunk_r + 1e-3;
        const double R2 = Rc * Rc;
        for (int base = 0; base < a.n_img; base += TILE) {
            // scan: keep the images within Rc of the chunk centre
            const int t = base + tid;
            bool keep = false;
            double4 rec = make_double4(0.0, 0.0, 0.0, 0.0);
            int ty = 0;
            if (t < a.n_img) {
                const double4 im = a.img[t];
                const double ex = im.x - cf0, ey = im.y - cf1, ez = im.z - cf2;
                const double cx = f11 * ex + f12 * ey + f13 * ez;
                const double cy = f22 * ey + f23 * ez;
                const double cz = f33 * ez;
                keep = cx * cx + cy * cy + cz * cz < R2;
                rec = make_double4(cx, cy, cz, im.w);
                ty = a.img_type[t] * nsite;
            }
            const unsigned long long m = __ballot(keep);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            __syncthreads();  // the previous tile's candidates are consumed
            if (lane == 0) s_wcnt[wave] = __popcll(m);
            __syncthreads();
            int woff[WG / 64 + 1];
            woff[0] = 0;
#pragma unroll
            for (int w = 0; w < WG / 64; ++w) woff[w + 1] = woff[w] + s_wcnt[w];
            if (keep) {
                s_c[woff[wave] + before] = rec;
                s_t[woff[wave] + before] = ty;
            }
            __syncthreads();
            n_cand_total += (unsigned long long)woff[WG / 64];
            if (wave_live) {
                // candidates of each 64-image group accumulate in Real, flushed to FP64
#pragma unroll
                for (int g = 0; g < WG / 64; ++g) {
                    Real plj = (Real)0, psr = (Real)0;
                    for (int j = woff[g]; j < woff[g + 1]; ++j) {
                        const double4 c = s_c[j];
                        const double ddx = c.x - prx, ddy = c.y - pry, ddz = c.z - prz;
                        const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        pair_term<ESM, Real>(r2, c.w, s_lj[s_t[j] + site], is_lj, r2lj, r2sr, alpha, inv_h,
                                             s_erfc, ovl, plj, psr);
                    }
                    acc_lj += (double)plj;
                    acc_sr += (double)psr;
                }
            }
        }
    } else {
        // fallback: every replicated atom, nearest image per pair
        for (int base = 0; base < a.n_atoms; base += TILE) {
            const int nt = min(TILE, a.n_atoms - base);
            __syncthreads();
            if (tid < nt) {
                s_c[tid] = a.atoms[base + tid];
                s_t[tid] = a.species[base + tid] * nsite;
            }
            __syncthreads();
            n_cand_total += (unsigned long long)nt;
            if (wave_live) {
                for (int g = 0; g < nt; g += 64) {
                    Real plj = (Real)0, psr = (Real)0;
                    const int ge = min(nt, g + 64);
                    for (int j = g; j < ge; ++j) {
                        const double4 at = s_c[j];
                        double ex = at.x - px, ey = at.y - py, ez = at.z - pz;
                        ex -= rint(ex);
                        ey -= rint(ey);
                        ez -= rint(ez);
                        const double cx = f11 * ex + f12 * ey + f13 * ez;
                        const double cy = f22 * ey + f23 * ez;
                        const double cz = f33 * ez;
                        const double r2 = cx * cx + cy * cy + cz * cz;
                        pair_term<ESM, Real>(r2, at.w, s_lj[s_t[j] + site], is_lj, r2lj, r2sr, alpha, inv_h,
                                             s_erfc, ovl, plj, psr);
                    }
                    acc_lj += (double)plj;
                    acc_sr += (double)psr;
                }
            }
        }
    }
    if (valid) {
        a.out_lj[pid] = ovl ? INFINITY : acc_lj;
        if constexpr (ES) a.out_sr[pid] = acc_sr;
    }
    if (a.cand_pairs != nullptr) {
        const unsigned long long m = __ballot(valid);
        if (lane == 0 && m != 0ull)
            atomicAdd(a.cand_pairs, (unsigned long long)__popcll(m) * n_cand_total);
    }
}

struct FArgs {
    int mode, nsite;
    long long n_slots;
    const double4* pos;        // MODE_WIDOM: per point slot (w = 0 marks an empty slot)
    int grid_n[3];
    double repd[3];
    unsigned site_q_mask;      // sites that carry a charge
    int Ma, Mb, Mc;            // Fourier box: ma in [0, Ma], mb in [-Mb, Mb], mc in [-Mc, Mc]
    const double2* A;          // dense box of wt * N_R * S(m), zeros off the kept set
    double* out_lr;
};

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Fourier part per point: sum_m Re(A_m e^{-i 2 pi m.u}), u = R o x (unit-cell
// fractional), by complex recurrences along each axis of the box (FP64).
__global__ __launch_bounds__(WG) void point_fourier_kernel(FArgs a) {
    __shared__ double2 s_a[FT];
    const int tid = threadIdx.x;
    const long long pid = (long long)blockIdx.x * WG + tid;
    bool valid = pid < a.n_slots;
    double x = 0.0, y = 0.0, z = 0.0;
    int site = 0;
    if (a.mode == MODE_WIDOM) {
        const double4 p = valid ? a.pos[pid] : make_double4(0.0, 0.0, 0.0, 0.0);
        valid = valid && p.w != 0.0;
        x = p.x; y = p.y; z = p.z;
        site = (int)(pid % a.nsite);
    } else if (valid) {
        const int n1 = a.grid_n[0], n2 = a.grid_n[1], n3 = a.grid_n[2];
        const long long i = pid % n1, j = (pid / n1) % n2, k = pid / ((long long)n1 * n2);
        x = grid_coord((double)i, n1, a.repd[0]);
        y = grid_coord((double)j, n2, a.repd[1]);
        z = grid_coord((double)k, n3, a.repd[2]);
    }
    const bool charged = valid && ((a.site_q_mask >> site) & 1u);
    double t0 = x * a.repd[0], t1 = y * a.repd[1], t2 = z * a.repd[2];
    t0 -= rint(t0);
    t1 -= rint(t1);
    t2 -= rint(t2);
    double s, c;
    sincospi(2.0 * t0, &s, &c);
    const double2 Ea = make_double2(c, -s);
    sincospi(2.0 * t1, &s, &c);
    const double2 Eb = make_double2(c, -s);
    sincospi(2.0 * t2, &s, &c);
    const double2 Ec = make_double2(c, -s);
    double2 Eb_neg = make_double2(1.0, 0.0), Ec_neg = make_double2(1.0, 0.0);
    const double2 Eb_c = make_double2(Eb.x, -Eb.y), Ec_c = make_double2(Ec.x, -Ec.y);
    for (int q = 0; q < a.Mb; ++q) Eb_neg = cmul(Eb_neg, Eb_c);
    for (int q = 0; q < a.Mc; ++q) Ec_neg = cmul(Ec_neg, Ec_c);

    const int rowlen = 2 * a.Mc + 1, nb = 2 * a.Mb + 1;
    const int n_rows = (a.Ma + 1) * nb;
    const int rows_per_tile = max(1, FT / rowlen);
    double2 Pa = make_double2(1.0, 0.0);
    double2 Pb = Eb_neg;
    double acc = 0.0;
    for (int r0 = 0; r0 < n_rows; r0 += rows_per_tile) {
        const int r1 = min(n_rows, r0 + rows_per_tile);
        const int n_ent = (r1 - r0) * rowlen;
        __syncthreads();
        for (int e = tid; e < n_ent; e += WG) s_a[e] = a.A[(long long)r0 * rowlen + e];
        __syncthreads();
        for (int r = r0; r < r1; ++r) {
            if (r > 0 && (r % nb) == 0) {  // next m_a
                Pa = cmul(Pa, Ea);
                Pb = cmul(Pa, Eb_neg);
            }
            double2 P = cmul(Pb, Ec_neg);
            const double2* row = s_a + (r - r0) * rowlen;
            for (int mc = 0; mc < rowlen; ++mc) {
                const double2 A = row[mc];
                acc += A.x * P.x - A.y * P.y;
                P = cmul(P, Ec);
            }
            Pb = cmul(Pb, Eb);
        }
    }
    if (valid) a.out_lr[pid] = charged ? acc : 0.0;
}

struct CArgs {
    int nsite, es;
    int n_blocks, chunks_per_block;
    long long n_per_block;
    int w0;
    double site_q[PM_MAX_SITES];
    double inv4pe, temperature;
    const double* lj;
    const double* sr;
    const double* lr;
    double* out_e;        // nullable: (E_vdw, E_es) per insertion
    double* partials;     // per insertion workgroup: (S, W_vdw, W_es)
};

// Per insertion: E_vdw = sum_s lj_s, E_es = sum_s q_s (sr_s / (4 pi eps0) + lr_s),
// Boltzmann weight and the workgroup's fixed-order partial sums.
__global__ __launch_bounds__(WG) void widom_combine_kernel(CArgs a) {
    const int w = a.w0 + blockIdx.x;
    const int b = w / a.chunks_per_block, c = w - b * a.chunks_per_block;
    const long long i = (long long)c * WG + threadIdx.x;
    const bool valid = i < a.n_per_block;
    const long long slot = (long long)blockIdx.x * WG + threadIdx.x;
    double vdw = 0.0, es = 0.0;
    if (valid) {
        for (int s = 0; s < a.nsite; ++s) {
            const long long p = slot * a.nsite + s;
            vdw += a.lj[p];
            if (a.es) es += a.site_q[s] * (a.sr[p] * a.inv4pe + a.lr[p]);
        }
    }
    const double wgt = exp(-(vdw + es) / a.temperature);
    const bool fin = !isinf(vdw) && !isinf(es);
    if (a.out_e != nullptr && valid) {
        const long long g = (long long)b * a.n_per_block + i;
        a.out_e[2 * g] = vdw;
        a.out_e[2 * g + 1] = es;
    }
    block_sum3(valid ? wgt : 0.0, (valid && fin) ? wgt * vdw : 0.0, (valid && fin) ? wgt * es : 0.0,
               a.partials + 3LL * w);
}

struct GCArgs {
    long long n;
    double q, inv4pe;
    int es;
    const double* lj;
    const double* sr;
    const double* lr;
    double* out_d;
    float* out_f;
};

__global__ void grid_combine_kernel(GCArgs a) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    double e = a.lj[p];
    if (a.es) e += a.q * (a.sr[p] * a.inv4pe + a.lr[p]);
    if (a.out_f != nullptr) a.out_f[p] = (float)e;
    else a.out_d[p] = e;
}

// Fixed-order per-block reduction of the workgroup partials: one workgroup
// per Widom block, strided partial sums then a shuffle tree.
__global__ __launch_bounds__(WG) void block_sums_kernel(const double* partials, int chunks, double* out) {
    const int b = blockIdx.x;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int c = threadIdx.x; c < chunks; c += WG) {
        const double* p = partials + 3LL * ((long long)b * chunks + c);
        t0 += p[0]; t1 += p[1]; t2 += p[2];
    }
    block_sum3(t0, t1, t2, out + 3LL * b);
}

}  // namespace pmk
