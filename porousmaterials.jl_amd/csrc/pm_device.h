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
constexpr double R2_OVERLAP = 0.1;  // src/vdw_energetics.jl:2

enum { MODE_WIDOM = 0, MODE_GRID = 2 };

struct KVec {
    double ma, mb, mc, are, aim;
    double wnr;  // wt * N_R (the factor of the structure factor in A)
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
        // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32)
        const uint64_t m0 = (uint64_t)0xD2511F53u * c[0], m1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(m0 >> 32), lo0 = (uint32_t)m0;
        uint32_t hi1 = (uint32_t)(m1 >> 32), lo1 = (uint32_t)m1;
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
    double c2f[9];                        // replicated-box c_to_f, column-major (forces)
    double repd[3];                       // replication factors (unit-cell phases)
    double r2_lj, r2_sr, alpha, inv4pe;
    Site site[PM_MAX_SITES];
    const double4* atoms;  // replicated atoms: fractional xyz + charge
    const int* species;
    const void* ljtab;     // [n_types][ns] of LJP<Real>
    const KVec* kv;
    const double* pose_xf; // n_items x ns x 3
    double* out_e;         // n_items x 3: (vdw, sr, lr)
    double* out_g;         // GRAD: n_items x ns x 3, dE/d(site Cartesian xyz) in K/A
};

// One wave per pose: the workgroup's four poses share LDS tiles of framework
// atoms (and of k-vectors), the wave's lanes stride over a tile, and FP64 lane
// partials are reduced by a fixed shuffle tree.  GRAD adds the gradient of the
// pose's energy with respect to every site's Cartesian position:
//   LJ      d/dx_s = 6 eps4 (2u^2 - u) / r^2 * d,   u = (sigma^2 / r^2)^3
//   real    d/dx_s = q_s q_a [erfc(a r) / r + 2a/sqrt(pi) e^{-a^2 r^2}] / r^2 * d / (4 pi eps0)
//   Fourier d/dx_s = q_s 2 pi C^T sum_k (-A_re sin th + A_im cos th) (m o rep),  th = 2 pi m.(rep o x)
// with d = x_atom - x_s (nearest image, Cartesian) and C the replicated box's c_to_f.
template <int NS, bool ES, typename Real, bool GRAD>
__global__ __launch_bounds__(WG) void pose_kernel(PoseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int ATOM_BYTES = TILE * (int)sizeof(double4);
    constexpr int K_BYTES = KTILE * (int)sizeof(KVec);
    constexpr int REGION = ATOM_BYTES > K_BYTES ? ATOM_BYTES : K_BYTES;
    double4* s_atoms = reinterpret_cast<double4*>(smem);
    KVec* s_k = reinterpret_cast<KVec*>(smem);
    int* s_type = reinterpret_cast<int*>(smem + REGION);
    LJP<Real>* s_lj = reinterpret_cast<LJP<Real>*>(smem + REGION + TILE * sizeof(int));

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const LJP<Real>* gtab = reinterpret_cast<const LJP<Real>*>(a.ljtab);
    for (int i = tid; i < a.n_types * NS; i += WG) s_lj[i] = gtab[i];

    const long long item = (long long)blockIdx.x * (WG / 64) + wave;
    const bool valid = item < a.n_items;
    const long long it = valid ? item : 0;
    double sx[NS], sy[NS], sz[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double* p = a.pose_xf + (it * NS + s) * 3;
        sx[s] = p[0]; sy[s] = p[1]; sz[s] = p[2];
    }

    double acc_lj[NS], acc_sr[NS];
    double gx[NS], gy[NS], gz[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        acc_lj[s] = 0.0; acc_sr[s] = 0.0;
        gx[s] = 0.0; gy[s] = 0.0; gz[s] = 0.0;
    }
    bool ovl = false;
    const double f11 = a.f11, f12 = a.f12, f13 = a.f13, f22 = a.f22, f23 = a.f23, f33 = a.f33;
    const double r2lj = a.r2_lj, r2sr = a.r2_sr, alpha = a.alpha;
    const double two_a_rpi = 2.0 * alpha * 0.56418958354775628;  // 2 alpha / sqrt(pi)
    for (int base = 0; base < a.n_atoms; base += TILE) {
        const int nt = min(TILE, a.n_atoms - base);
        __syncthreads();
        if (tid < nt) {
            s_atoms[tid] = a.atoms[base + tid];
            s_type[tid] = a.species[base + tid];
        }
        __syncthreads();
        if (!valid) continue;
        for (int j = lane; j < nt; j += 64) {
            const double4 at = s_atoms[j];
            const int t = s_type[j];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double dx = at.x - sx[s], dy = at.y - sy[s], dz = at.z - sz[s];
                dx -= rint(dx);
                dy -= rint(dy);
                dz -= rint(dz);
                const double cx = f11 * dx + f12 * dy + f13 * dz;
                const double cy = f22 * dy + f23 * dz;
                const double cz = f33 * dz;
                const double r2 = cx * cx + cy * cy + cz * cz;
                const LJP<Real> p = s_lj[t * NS + s];
                const bool in = r2 < r2lj;
                ovl |= in & (r2 < R2_OVERLAP) & (a.site[s].lj != 0);
                const Real e = PairMath<Real>::lj(r2, p.sig2, p.eps4);
                acc_lj[s] += in ? (double)e : 0.0;
                if (GRAD) {
                    // 6 eps4 (2u^2 - u) / r^2
                    const double rr = 1.0 / r2;
                    const double sg = (double)p.sig2 * rr;
                    const double u = sg * sg * sg;
                    const double f = in ? 6.0 * (double)p.eps4 * u * (2.0 * u - 1.0) * rr : 0.0;
                    gx[s] += f * cx;
                    gy[s] += f * cy;
                    gz[s] += f * cz;
                }
                if (ES) {
                    const bool ins = (r2 < r2sr) & (at.w != 0.0);
                    const Real c = PairMath<Real>::coul(r2, at.w, alpha);
                    acc_sr[s] += ins ? (double)c : 0.0;
                    if (GRAD) {
                        const double r = sqrt(r2);
                        const double f = ins ? at.w * (erfc(alpha * r) / r + two_a_rpi * exp(-alpha * alpha * r2)) /
                                                   r2 * a.site[s].q * a.inv4pe
                                             : 0.0;
                        gx[s] += f * cx;
                        gy[s] += f * cy;
                        gz[s] += f * cz;
                    }
                }
            }
        }
    }

    double acc_lr[NS];
    double hx[NS], hy[NS], hz[NS];  // Fourier gradient in (m o rep) space
#pragma unroll
    for (int s = 0; s < NS; ++s) { acc_lr[s] = 0.0; hx[s] = 0.0; hy[s] = 0.0; hz[s] = 0.0; }
    if (ES) {
        double ux[NS], uy[NS], uz[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            ux[s] = sx[s] * a.repd[0];
            uy[s] = sy[s] * a.repd[1];
            uz[s] = sz[s] * a.repd[2];
        }
        for (int base = 0; base < a.nk; base += KTILE) {
            const int nt = min(KTILE, a.nk - base);
            __syncthreads();
            if (tid < nt) s_k[tid] = a.kv[base + tid];
            __syncthreads();
            if (!valid) continue;
            for (int k = lane; k < nt; k += 64) {
                const KVec kv = s_k[k];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    if (a.site[s].q != 0.0) {
                        double t = kv.ma * ux[s] + kv.mb * uy[s] + kv.mc * uz[s];
                        t -= rint(t);
                        double sn, cs;
                        sincospi(2.0 * t, &sn, &cs);
                        acc_lr[s] += kv.are * cs + kv.aim * sn;
                        if (GRAD) {
                            const double d = kv.aim * cs - kv.are * sn;
                            hx[s] += d * kv.ma;
                            hy[s] += d * kv.mb;
                            hz[s] += d * kv.mc;
                        }
                    }
                }
            }
        }
    }

    // fixed-order reductions over the wave's lanes
    double vdw = 0.0, sr = 0.0, lr = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        vdw += wave_sum(acc_lj[s]);
        if (ES) {
            sr += a.site[s].q * wave_sum(acc_sr[s]);
            lr += a.site[s].q * wave_sum(acc_lr[s]);
        }
    }
    ovl = __ballot(ovl) != 0ull;
    if (ovl) vdw = INFINITY;
    if (GRAD) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            gx[s] = wave_sum(gx[s]);
            gy[s] = wave_sum(gy[s]);
            gz[s] = wave_sum(gz[s]);
            if (ES) {
                hx[s] = wave_sum(hx[s]);
                hy[s] = wave_sum(hy[s]);
                hz[s] = wave_sum(hz[s]);
            }
        }
    }
    if (valid && lane == 0) {
        a.out_e[3 * item + 0] = vdw;
        a.out_e[3 * item + 1] = sr * a.inv4pe;
        a.out_e[3 * item + 2] = lr;
        if (GRAD) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // Fourier: 2 pi q_s C^T (h o rep)
                const double q2p = 2.0 * M_PI * a.site[s].q;
                const double vx = hx[s] * a.repd[0] * q2p, vy = hy[s] * a.repd[1] * q2p, vz = hz[s] * a.repd[2] * q2p;
                const double* C = a.c2f;
                double* o = a.out_g + (item * NS + s) * 3;
                o[0] = gx[s] + C[0] * vx + C[1] * vy + C[2] * vz;
                o[1] = gy[s] + C[3] * vx + C[4] * vy + C[5] * vz;
                o[2] = gz[s] + C[6] * vx + C[7] * vy + C[8] * vz;
            }
        }
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
        kv.wnr = wt[k] * nr;
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

}  // namespace pmk
