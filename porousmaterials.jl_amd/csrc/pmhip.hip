// pmhip.hip -- gfx950 (CDNA4, wave64) kernels and C ABI for guest-host
// Lennard-Jones + Ewald energies, Widom insertions and energy grids.
//
// Numerical contract (restating the reference, PorousMaterials.jl v0.4.2):
//   LJ pair        src/vdw_energetics.jl:19-22    4 eps s^3 (s^3 - 1), s = sigma^2/r^2
//   pair loop      src/vdw_energetics.jl:25-47    nearest image in fractional space,
//                                                 strict r^2 < rc^2, overlap r^2 < 0.1 -> +Inf
//   real space     src/electrostatics.jl:377-386  q erfc(alpha r) / (4 pi eps0 r), r < rc
//   Fourier        src/electrostatics.jl:354-372  sum_k wt_k sum_c q_c cos(k.(x_c - x)),
//                                                 evaluated here from the unit-cell
//                                                 structure factor S(k) (identical sum)
//   Widom          src/henry.jl:212-243           per-block sum e^{-E/T}, E e^{-E/T}
//   energy grid    src/grid.jl:313-355            point (i,j,k)/(n-1) ./ rep
//
// Layout in HBM: replicated framework atoms as double4 (x, y, z fractional in the
// replicated box, charge) plus an int32 compact species index; k-vectors as
// KVec {m_a, m_b, m_c, Re A, Im A} with A = wt * N_R * S(m) in FP64.
//
// Execution: one work-item per insertion / pose / grid point, 256 work-items per
// workgroup; the workgroup streams framework atoms through LDS in tiles of 256
// (coalesced double4 loads), every lane reads the same LDS record (broadcast), so
// LDS traffic is one record per atom per wave and the pair arithmetic is VALU-bound.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pmhip.h"

#define PM_VERSION "pmhip 0.1.0 (gfx950)"

namespace {

constexpr int WG = 256;      // work-items per workgroup (4 waves)
constexpr int TILE = 256;    // framework atoms per LDS tile
constexpr int KTILE = 128;   // k-vectors per LDS tile
constexpr double R2_OVERLAP = 0.1;  // src/vdw_energetics.jl:2

enum { MODE_WIDOM = 0, MODE_POSES = 1, MODE_GRID = 2 };

struct KVec {
    double ma, mb, mc, are, aim, pad;
};

struct Site {
    double ox, oy, oz;  // Cartesian offset from the centre of mass (template)
    double q;           // point charge (e), 0 if none
    int lj;             // 1 if an LJ sphere
    int pad;
};

struct KArgs {
    int mode, ns, n_atoms, n_types, nk, rotate;
    int n_blocks, chunks_per_block;
    long long n_per_block, n_items;
    int grid_n[3];
    int pad0;
    double f11, f12, f13, f22, f23, f33;  // replicated-box f_to_c (upper triangular)
    double c2f[9];                        // its inverse, column-major
    double repd[3];                       // replication factors (unit-cell phases)
    double r2_lj, r2_sr, alpha, inv4pe, temperature;
    unsigned long long seed;
    Site site[PM_MAX_SITES];
    const double4* atoms;
    const int* species;
    const void* ljtab;      // [n_types][ns] of (sigma^2, 4 eps) in Real
    const KVec* kv;
    const double* pose_xf;  // MODE_POSES: n_items x ns x 3
    double* out_e;          // per-item outputs (mode dependent)
    float* out_f;           // MODE_GRID, FP32 grid
    double* partials;       // MODE_WIDOM: per-workgroup (S, Wvdw, Wes)
};

template <typename Real>
struct LJP {
    Real sig2, eps4;
};

// ---------------------------------------------------------------- RNG ----
// Philox4x32-10 (Salmon et al., SC'11); same stream layout as the oracle:
// counter (g_lo, g_hi, w, 0x57494430), key (seed_lo, seed_hi), w = 0, 1, 2.
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

__device__ inline void insertion_uniforms(unsigned long long seed, unsigned long long g,
                                          double u[6]) {
#pragma unroll
    for (uint32_t w = 0; w < 3; ++w) {
        uint32_t c[4] = {(uint32_t)g, (uint32_t)(g >> 32), w, 0x57494430u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        u[2 * w] = u53(c[0], c[1]);
        u[2 * w + 1] = u53(c[2], c[3]);
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
    float y = p - x * x;  // exponent, natural log units
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

// ------------------------------------------------------------- kernel ----
template <int NS, bool ES, typename Real>
__global__ __launch_bounds__(WG) void energy_kernel(KArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    // [ max(atom tile, k tile) | types | LJ table ]
    constexpr int ATOM_BYTES = TILE * (int)sizeof(double4);
    constexpr int K_BYTES = KTILE * (int)sizeof(KVec);
    constexpr int REGION = ATOM_BYTES > K_BYTES ? ATOM_BYTES : K_BYTES;
    double4* s_atoms = reinterpret_cast<double4*>(smem);
    KVec* s_k = reinterpret_cast<KVec*>(smem);
    int* s_type = reinterpret_cast<int*>(smem + REGION);
    LJP<Real>* s_lj = reinterpret_cast<LJP<Real>*>(smem + REGION + TILE * sizeof(int));

    const int tid = threadIdx.x;
    const int ntab = a.n_types * NS;
    const LJP<Real>* gtab = reinterpret_cast<const LJP<Real>*>(a.ljtab);
    for (int i = tid; i < ntab; i += WG) s_lj[i] = gtab[i];

    // ---- work item and its site positions (replicated-box fractional) ----
    long long item;
    bool valid;
    if (a.mode == MODE_WIDOM) {
        int b = blockIdx.x / a.chunks_per_block;
        int c = blockIdx.x - b * a.chunks_per_block;
        long long i = (long long)c * WG + tid;
        valid = i < a.n_per_block;
        item = (long long)b * a.n_per_block + i;
    } else {
        item = (long long)blockIdx.x * WG + tid;
        valid = item < a.n_items;
    }

    double sx[NS], sy[NS], sz[NS];
    if (a.mode == MODE_WIDOM) {
        double u[6];
        insertion_uniforms(a.seed, (unsigned long long)item, u);
        double R[9];  // column-major rotation, Arvo: -H * Rz (src/molecule.jl:233-244)
        if (a.rotate) {
            double s1, c1, s2, c2;
            sincospi(2.0 * u[3], &s1, &c1);
            sincospi(2.0 * u[4], &s2, &c2);
            double sq = sqrt(u[5]);
            double v0 = c2 * sq, v1 = s2 * sq, v2 = sqrt(1.0 - u[5]);
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
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double y0 = a.site[s].ox, y1 = a.site[s].oy, y2 = a.site[s].oz;
            if (a.rotate) {
                double t0 = R[0] * y0 + R[3] * y1 + R[6] * y2;
                double t1 = R[1] * y0 + R[4] * y1 + R[7] * y2;
                double t2 = R[2] * y0 + R[5] * y1 + R[8] * y2;
                y0 = t0; y1 = t1; y2 = t2;
            }
            sx[s] = u[0] + (a.c2f[0] * y0 + a.c2f[3] * y1 + a.c2f[6] * y2);
            sy[s] = u[1] + (a.c2f[1] * y0 + a.c2f[4] * y1 + a.c2f[7] * y2);
            sz[s] = u[2] + (a.c2f[2] * y0 + a.c2f[5] * y1 + a.c2f[8] * y2);
        }
    } else if (a.mode == MODE_POSES) {
        long long it = valid ? item : 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double* p = a.pose_xf + (it * NS + s) * 3;
            sx[s] = p[0]; sy[s] = p[1]; sz[s] = p[2];
        }
    } else {
        long long it = valid ? item : 0;
        int n1 = a.grid_n[0], n2 = a.grid_n[1], n3 = a.grid_n[2];
        long long i = it % n1, j = (it / n1) % n2, k = it / ((long long)n1 * n2);
        double x = (n1 > 1 ? (double)i / (double)(n1 - 1) : 0.0) / a.repd[0];
        double y = (n2 > 1 ? (double)j / (double)(n2 - 1) : 0.0) / a.repd[1];
        double z = (n3 > 1 ? (double)k / (double)(n3 - 1) : 0.0) / a.repd[2];
#pragma unroll
        for (int s = 0; s < NS; ++s) { sx[s] = x; sy[s] = y; sz[s] = z; }
    }

    // ---- pair loop over framework atoms (LJ + real-space Ewald) ----
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

    // ---- Fourier part from the unit-cell structure factor (FP64) ----
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
    sr *= a.inv4pe;

    if (a.mode == MODE_WIDOM) {
        double es = sr + lr;
        double w = exp(-(vdw + es) / a.temperature);
        bool fin = !isinf(vdw) && !isinf(es);
        double v0 = valid ? w : 0.0;
        double v1 = (valid && fin) ? w * vdw : 0.0;
        double v2 = (valid && fin) ? w * es : 0.0;
        if (a.out_e != nullptr && valid) {
            a.out_e[2 * item] = vdw;
            a.out_e[2 * item + 1] = es;
        }
        __shared__ double red[3][WG / 64];
        v0 = wave_sum(v0);
        v1 = wave_sum(v1);
        v2 = wave_sum(v2);
        int wave = tid >> 6, lane = tid & 63;
        if (lane == 0) { red[0][wave] = v0; red[1][wave] = v1; red[2][wave] = v2; }
        __syncthreads();
        if (tid == 0) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            for (int w2 = 0; w2 < WG / 64; ++w2) { t0 += red[0][w2]; t1 += red[1][w2]; t2 += red[2][w2]; }
            a.partials[3 * blockIdx.x + 0] = t0;
            a.partials[3 * blockIdx.x + 1] = t1;
            a.partials[3 * blockIdx.x + 2] = t2;
        }
    } else if (a.mode == MODE_POSES) {
        if (valid) {
            a.out_e[3 * item + 0] = vdw;
            a.out_e[3 * item + 1] = sr;
            a.out_e[3 * item + 2] = lr;
        }
    } else {
        if (valid) {
            double e = vdw + sr + lr;
            if (a.out_f != nullptr) a.out_f[item] = (float)e;
            else a.out_e[item] = e;
        }
    }
}

// Fixed-order per-block reduction of the workgroup partials (deterministic).
__global__ void block_sums_kernel(const double* partials, int chunks, int n_blocks, double* out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_blocks) return;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double* p = partials + 3LL * ((long long)b * chunks + c);
        t0 += p[0]; t1 += p[1]; t2 += p[2];
    }
    out[3 * b] = t0; out[3 * b + 1] = t1; out[3 * b + 2] = t2;
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

// ------------------------------------------------------------ host side ----
thread_local std::string g_err;
std::mutex g_mu;

struct Ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, last_pairs = 0.0;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // host staging for the per-call LJ table: outlives the asynchronous copy
    std::vector<unsigned char> host_tab;
};
Ctx g_ctx;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define PM_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(e_ == hipErrorOutOfMemory ? PM_ERR_OOM : PM_ERR_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

int ensure_scratch(size_t bytes) {
    if (bytes <= g_ctx.scratch_bytes) return PM_OK;
    if (g_ctx.scratch) {
        PM_HIP(hipStreamSynchronize(g_ctx.stream));
        PM_HIP(hipFree(g_ctx.scratch));
        g_ctx.scratch = nullptr;
        g_ctx.scratch_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    PM_HIP(hipMalloc(&g_ctx.scratch, want));
    g_ctx.scratch_bytes = want;
    return PM_OK;
}

int require_init() {
    if (g_ctx.device < 0) return fail(PM_ERR_STATE, "pm_init has not been called");
    return PM_OK;
}

void inv3(const double m[9], double out[9]) {
    double d = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
               m[6] * (m[1] * m[5] - m[4] * m[2]);
    out[0] = (m[4] * m[8] - m[7] * m[5]) / d;
    out[3] = -(m[3] * m[8] - m[6] * m[5]) / d;
    out[6] = (m[3] * m[7] - m[6] * m[4]) / d;
    out[1] = -(m[1] * m[8] - m[7] * m[2]) / d;
    out[4] = (m[0] * m[8] - m[6] * m[2]) / d;
    out[7] = -(m[0] * m[7] - m[6] * m[1]) / d;
    out[2] = (m[1] * m[5] - m[4] * m[2]) / d;
    out[5] = -(m[0] * m[5] - m[3] * m[2]) / d;
    out[8] = (m[0] * m[4] - m[3] * m[1]) / d;
}

}  // namespace

struct pm_framework {
    double F[9];      // replicated-box f_to_c
    double c2f[9];
    int32_t rep[3];
    int64_t n_unit = 0, n_rep = 0, n_unit_q = 0;
    std::vector<int32_t> species_global;  // compact -> force-field species
    double net_charge = 0.0;
    bool charged = false;
    double4* d_atoms = nullptr;
    int32_t* d_species = nullptr;
    double4* d_unit_q = nullptr;
};

struct pm_forcefield {
    int32_t n = 0;
    std::vector<double> eps, sig2;
    double r2_cutoff = 0.0;
};

struct pm_ewald {
    double alpha = 0.0, sr_cutoff = 0.0;
    int64_t nk = 0;
    KVec* d_k = nullptr;
};

struct pm_probe {
    int ns = 0;
    int rotate = 0;
    Site site[PM_MAX_SITES];
    int32_t lj_species[PM_MAX_SITES];
    bool charged = false;
};

namespace {

// Fill kernel arguments common to every launch.
void fill_box(KArgs& a, const pm_framework* fw) {
    a.f11 = fw->F[0]; a.f12 = fw->F[3]; a.f13 = fw->F[6];
    a.f22 = fw->F[4]; a.f23 = fw->F[7]; a.f33 = fw->F[8];
    for (int i = 0; i < 9; ++i) a.c2f[i] = fw->c2f[i];
    for (int i = 0; i < 3; ++i) a.repd[i] = (double)fw->rep[i];
    a.n_atoms = (int)fw->n_rep;
    a.n_types = (int)fw->species_global.size();
    a.atoms = fw->d_atoms;
    a.species = fw->d_species;
    a.inv4pe = 1.0 / (4.0 * M_PI * (8.854187817e-12 / (1.602176565e-19 * 1.602176565e-19) /
                                    1.0e10 * 1.3806488e-23));
}

// LJ table [n_types][ns] of (sigma^2, 4 eps) for the framework's compact
// species against each probe site, in the precision of the pair maths.
template <typename Real>
void build_ljtab(const pm_framework* fw, const pm_forcefield* ff, int ns, const Site* site,
                 const int32_t* site_species, std::vector<unsigned char>& out) {
    int nt = (int)fw->species_global.size();
    out.resize(sizeof(LJP<Real>) * (size_t)(nt > 0 ? nt : 1) * ns);
    LJP<Real>* t = reinterpret_cast<LJP<Real>*>(out.data());
    for (int i = 0; i < nt; ++i)
        for (int s = 0; s < ns; ++s) {
            LJP<Real> p;
            p.sig2 = (Real)1.0;
            p.eps4 = (Real)0.0;
            if (site[s].lj && ff != nullptr) {
                size_t idx = (size_t)fw->species_global[i] * ff->n + site_species[s];
                p.sig2 = (Real)ff->sig2[idx];
                p.eps4 = (Real)(4.0 * ff->eps[idx]);
            }
            t[i * ns + s] = p;
        }
}

size_t lds_bytes(int n_types, int ns, bool fp64) {
    size_t region = (size_t)TILE * sizeof(double4);
    size_t kb = (size_t)KTILE * sizeof(KVec);
    if (kb > region) region = kb;
    size_t ljb = (size_t)(n_types > 0 ? n_types : 1) * ns * (fp64 ? 16 : 8);
    return region + TILE * sizeof(int) + ljb;
}

template <int NS, bool ES, typename Real>
void launch_t(const KArgs& a, int n_wg, size_t shm) {
    hipLaunchKernelGGL((energy_kernel<NS, ES, Real>), dim3(n_wg), dim3(WG), shm, g_ctx.stream, a);
}

template <bool ES, typename Real>
int launch_ns(const KArgs& a, int n_wg, size_t shm) {
    switch (a.ns) {
        case 1: launch_t<1, ES, Real>(a, n_wg, shm); break;
        case 2: launch_t<2, ES, Real>(a, n_wg, shm); break;
        case 3: launch_t<3, ES, Real>(a, n_wg, shm); break;
        case 4: launch_t<4, ES, Real>(a, n_wg, shm); break;
        case 5: launch_t<5, ES, Real>(a, n_wg, shm); break;
        case 6: launch_t<6, ES, Real>(a, n_wg, shm); break;
        case 8: launch_t<8, ES, Real>(a, n_wg, shm); break;
        default: return fail(PM_ERR_ARG, "unsupported site count " + std::to_string(a.ns));
    }
    return PM_OK;
}

// Launch the energy kernel with timing events around it.
int launch_energy(const KArgs& a, bool es, int precision, int n_wg) {
    bool fp64 = precision == PM_FP64;
    size_t shm = lds_bytes(a.n_types, a.ns, fp64);
    if (n_wg <= 0) return PM_OK;
    PM_HIP(hipEventRecord(g_ctx.ev0, g_ctx.stream));
    int rc;
    if (fp64) rc = es ? launch_ns<true, double>(a, n_wg, shm) : launch_ns<false, double>(a, n_wg, shm);
    else rc = es ? launch_ns<true, float>(a, n_wg, shm) : launch_ns<false, float>(a, n_wg, shm);
    if (rc != PM_OK) return rc;
    PM_HIP(hipGetLastError());
    PM_HIP(hipEventRecord(g_ctx.ev1, g_ctx.stream));
    return PM_OK;
}

int finish_timing(double pairs) {
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    float ms = 0.f;
    PM_HIP(hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1));
    g_ctx.last_ms = ms;
    g_ctx.last_pairs = pairs;
    return PM_OK;
}

// Validate a handle's species against a force field.
int check_species(const pm_framework* fw, const pm_forcefield* ff) {
    for (int32_t s : fw->species_global)
        if (s < 0 || s >= ff->n)
            return fail(PM_ERR_ARG, "framework species index " + std::to_string(s) +
                                        " outside the force field's " + std::to_string(ff->n) +
                                        " species");
    return PM_OK;
}

}  // namespace

// ================================================================ C ABI ====
extern "C" {

const char* pm_version(void) { return PM_VERSION; }

const char* pm_last_error(void) { return g_err.c_str(); }

int pm_init(int32_t device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx.device == device) return PM_OK;
    if (g_ctx.device >= 0) return fail(PM_ERR_STATE, "pm_init: already initialised on another device");
    int n = 0;
    PM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(PM_ERR_ARG, "pm_init: device " + std::to_string(device) + " not present (" +
                                    std::to_string(n) + " visible)");
    PM_HIP(hipSetDevice(device));
    PM_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    PM_HIP(hipEventCreate(&g_ctx.ev0));
    PM_HIP(hipEventCreate(&g_ctx.ev1));
    g_ctx.device = device;
    return PM_OK;
}

int pm_framework_create(const double f_to_c[9], const int32_t rep[3], int64_t n_atoms,
                        const double* xf, const int32_t* species, const double* charge,
                        pm_framework** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!f_to_c || !rep || !out || n_atoms < 0 || (n_atoms > 0 && (!xf || !species)))
        return fail(PM_ERR_ARG, "pm_framework_create: null pointer or negative atom count");
    if (f_to_c[1] != 0.0 || f_to_c[2] != 0.0 || f_to_c[5] != 0.0)
        return fail(PM_ERR_ARG, "pm_framework_create: f_to_c must be upper-triangular (a along x)");
    if (!(f_to_c[0] > 0.0 && f_to_c[4] > 0.0 && f_to_c[8] > 0.0))
        return fail(PM_ERR_ARG, "pm_framework_create: f_to_c must have a positive diagonal");
    for (int i = 0; i < 3; ++i)
        if (rep[i] < 1 || rep[i] > 1024)
            return fail(PM_ERR_ARG, "pm_framework_create: replication factors must be in [1, 1024]");
    int64_t nr = (int64_t)rep[0] * rep[1] * rep[2];
    if (n_atoms * nr > (int64_t)INT32_MAX - TILE)
        return fail(PM_ERR_ARG, "pm_framework_create: replicated atom count exceeds 2^31");
    auto* fw = new pm_framework();
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) fw->F[r + 3 * c] = f_to_c[r + 3 * c] * rep[c];
    inv3(fw->F, fw->c2f);
    for (int i = 0; i < 3; ++i) fw->rep[i] = rep[i];
    fw->n_unit = n_atoms;
    fw->n_rep = n_atoms * nr;
    // compact species
    std::vector<int32_t> local(n_atoms);
    for (int64_t i = 0; i < n_atoms; ++i) {
        int32_t s = species[i];
        if (s < 0) { delete fw; return fail(PM_ERR_ARG, "pm_framework_create: negative species index"); }
        int32_t found = -1;
        for (size_t k = 0; k < fw->species_global.size(); ++k)
            if (fw->species_global[k] == s) { found = (int32_t)k; break; }
        if (found < 0) {
            if ((int)fw->species_global.size() >= PM_MAX_FW_SPECIES) {
                delete fw;
                return fail(PM_ERR_ARG, "pm_framework_create: more than PM_MAX_FW_SPECIES species");
            }
            found = (int32_t)fw->species_global.size();
            fw->species_global.push_back(s);
        }
        local[i] = found;
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(xf[3 * i + d])) {
                delete fw;
                return fail(PM_ERR_ARG, "pm_framework_create: non-finite coordinate");
            }
        if (charge && !std::isfinite(charge[i])) {
            delete fw;
            return fail(PM_ERR_ARG, "pm_framework_create: non-finite charge");
        }
    }
    std::vector<double4> atoms((size_t)fw->n_rep);
    std::vector<int32_t> sp((size_t)fw->n_rep);
    std::vector<double4> uq;
    int64_t k = 0;
    for (int ra = 0; ra < rep[0]; ++ra)
        for (int rb = 0; rb < rep[1]; ++rb)
            for (int rc = 0; rc < rep[2]; ++rc)
                for (int64_t i = 0; i < n_atoms; ++i, ++k) {
                    double q = charge ? charge[i] : 0.0;
                    atoms[k] = make_double4((xf[3 * i] + 1.0 * ra) / rep[0],
                                            (xf[3 * i + 1] + 1.0 * rb) / rep[1],
                                            (xf[3 * i + 2] + 1.0 * rc) / rep[2], q);
                    sp[k] = local[i];
                }
    for (int64_t i = 0; i < n_atoms; ++i) {
        double q = charge ? charge[i] : 0.0;
        fw->net_charge += q;
        if (q != 0.0) uq.push_back(make_double4(xf[3 * i], xf[3 * i + 1], xf[3 * i + 2], q));
    }
    fw->n_unit_q = (int64_t)uq.size();
    fw->charged = fw->n_unit_q > 0;
    auto cleanup = [&]() {
        hipFree(fw->d_atoms); hipFree(fw->d_species); hipFree(fw->d_unit_q);
        delete fw;
    };
    size_t na = (size_t)(fw->n_rep > 0 ? fw->n_rep : 1);
    if (hipMalloc(&fw->d_atoms, na * sizeof(double4)) != hipSuccess ||
        hipMalloc(&fw->d_species, na * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&fw->d_unit_q, (uq.empty() ? 1 : uq.size()) * sizeof(double4)) != hipSuccess) {
        cleanup();
        return fail(PM_ERR_OOM, "pm_framework_create: device allocation failed");
    }
    if (fw->n_rep > 0) {
        if (hipMemcpy(fw->d_atoms, atoms.data(), fw->n_rep * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(fw->d_species, sp.data(), fw->n_rep * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            return fail(PM_ERR_DEVICE, "pm_framework_create: host-to-device copy failed");
        }
    }
    if (!uq.empty() &&
        hipMemcpy(fw->d_unit_q, uq.data(), uq.size() * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(PM_ERR_DEVICE, "pm_framework_create: host-to-device copy failed");
    }
    *out = fw;
    return PM_OK;
}

int pm_framework_destroy(pm_framework* fw) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!fw) return PM_OK;
    hipFree(fw->d_atoms);
    hipFree(fw->d_species);
    hipFree(fw->d_unit_q);
    delete fw;
    return PM_OK;
}

int pm_framework_info(const pm_framework* fw, int64_t* n_replicated, int32_t* n_species,
                      double* net_charge) {
    if (!fw) return fail(PM_ERR_ARG, "pm_framework_info: null handle");
    if (n_replicated) *n_replicated = fw->n_rep;
    if (n_species) *n_species = (int32_t)fw->species_global.size();
    if (net_charge) *net_charge = fw->net_charge;
    return PM_OK;
}

int pm_forcefield_create(int32_t n_species, const double* epsilon, const double* sigma2,
                         double r2_cutoff, pm_forcefield** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out || n_species <= 0 || !epsilon || !sigma2)
        return fail(PM_ERR_ARG, "pm_forcefield_create: empty table or null pointer");
    if (!(r2_cutoff > 0.0) || !std::isfinite(r2_cutoff))
        return fail(PM_ERR_ARG, "pm_forcefield_create: r2_cutoff must be positive and finite");
    auto* ff = new pm_forcefield();
    ff->n = n_species;
    size_t nn = (size_t)n_species * n_species;
    ff->eps.assign(epsilon, epsilon + nn);
    ff->sig2.assign(sigma2, sigma2 + nn);
    for (size_t i = 0; i < nn; ++i)
        if (!std::isfinite(ff->eps[i]) || !std::isfinite(ff->sig2[i]) || ff->sig2[i] <= 0.0) {
            delete ff;
            return fail(PM_ERR_ARG, "pm_forcefield_create: non-finite epsilon or non-positive sigma^2");
        }
    ff->r2_cutoff = r2_cutoff;
    *out = ff;
    return PM_OK;
}

int pm_forcefield_destroy(pm_forcefield* ff) {
    delete ff;
    return PM_OK;
}

int pm_ewald_create(const pm_framework* fw, double alpha, double sr_cutoff, int64_t n_kvecs,
                    const int32_t* kidx, const double* wt, pm_ewald** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !out || n_kvecs < 0 || (n_kvecs > 0 && (!kidx || !wt)))
        return fail(PM_ERR_ARG, "pm_ewald_create: null pointer or negative k-vector count");
    if (!(alpha > 0.0) || !(sr_cutoff > 0.0) || !std::isfinite(alpha) || !std::isfinite(sr_cutoff))
        return fail(PM_ERR_ARG, "pm_ewald_create: alpha and sr_cutoff must be positive and finite");
    std::vector<int32_t> m;
    std::vector<double> w;
    for (int64_t k = 0; k < n_kvecs; ++k) {
        int32_t na = kidx[3 * k], nb = kidx[3 * k + 1], nc = kidx[3 * k + 2];
        if (!std::isfinite(wt[k])) return fail(PM_ERR_ARG, "pm_ewald_create: non-finite weight");
        if (na % fw->rep[0] || nb % fw->rep[1] || nc % fw->rep[2]) continue;
        m.push_back(na / fw->rep[0]);
        m.push_back(nb / fw->rep[1]);
        m.push_back(nc / fw->rep[2]);
        w.push_back(wt[k]);
    }
    auto* ew = new pm_ewald();
    ew->alpha = alpha;
    ew->sr_cutoff = sr_cutoff;
    ew->nk = (int64_t)w.size();
    if (ew->nk > 0) {
        int* d_m = nullptr;
        double* d_w = nullptr;
        if (hipMalloc(&ew->d_k, ew->nk * sizeof(KVec)) != hipSuccess ||
            hipMalloc(&d_m, m.size() * sizeof(int)) != hipSuccess ||
            hipMalloc(&d_w, w.size() * sizeof(double)) != hipSuccess) {
            hipFree(ew->d_k); hipFree(d_m); hipFree(d_w);
            delete ew;
            return fail(PM_ERR_OOM, "pm_ewald_create: device allocation failed");
        }
        hipMemcpyAsync(d_m, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, g_ctx.stream);
        hipMemcpyAsync(d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, g_ctx.stream);
        double nr = (double)fw->rep[0] * fw->rep[1] * fw->rep[2];
        hipLaunchKernelGGL(structure_factor_kernel, dim3((unsigned)ew->nk), dim3(64), 0, g_ctx.stream,
                           fw->d_unit_q, (long long)fw->n_unit_q, d_m, d_w, nr, ew->d_k);
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipStreamSynchronize(g_ctx.stream);
        hipFree(d_m);
        hipFree(d_w);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            hipFree(ew->d_k);
            delete ew;
            return fail(PM_ERR_DEVICE, std::string("pm_ewald_create: structure factor kernel: ") +
                                           hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        }
    }
    *out = ew;
    return PM_OK;
}

int pm_ewald_destroy(pm_ewald* ew) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ew) return PM_OK;
    hipFree(ew->d_k);
    delete ew;
    return PM_OK;
}

int pm_ewald_info(const pm_ewald* ew, int64_t* n_kept) {
    if (!ew) return fail(PM_ERR_ARG, "pm_ewald_info: null handle");
    if (n_kept) *n_kept = ew->nk;
    return PM_OK;
}

int pm_probe_create(int32_t n_lj, const double* lj_xyz, const int32_t* lj_species, int32_t n_q,
                    const double* q_xyz, const double* q, pm_probe** out) {
    if (!out || n_lj < 0 || n_q < 0 || (n_lj > 0 && (!lj_xyz || !lj_species)) ||
        (n_q > 0 && (!q_xyz || !q)))
        return fail(PM_ERR_ARG, "pm_probe_create: null pointer or negative count");
    if (n_lj + n_q == 0) return fail(PM_ERR_ARG, "pm_probe_create: probe has no sites");
    auto* pr = new pm_probe();
    pr->rotate = (n_lj + n_q) > 1;  // needs_rotations, src/mc_helpers.jl:8
    std::vector<bool> used(n_q, false);
    for (int32_t i = 0; i < n_lj; ++i) {
        if (lj_species[i] < 0) { delete pr; return fail(PM_ERR_ARG, "pm_probe_create: negative species"); }
        if (pr->ns >= PM_MAX_SITES) { delete pr; return fail(PM_ERR_ARG, "pm_probe_create: too many sites"); }
        Site s{};
        s.ox = lj_xyz[3 * i]; s.oy = lj_xyz[3 * i + 1]; s.oz = lj_xyz[3 * i + 2];
        s.lj = 1;
        s.q = 0.0;
        for (int32_t c = 0; c < n_q; ++c)
            if (!used[c] && q_xyz[3 * c] == s.ox && q_xyz[3 * c + 1] == s.oy && q_xyz[3 * c + 2] == s.oz) {
                used[c] = true;
                s.q = q[c];
                break;
            }
        pr->lj_species[pr->ns] = lj_species[i];
        pr->site[pr->ns++] = s;
    }
    for (int32_t c = 0; c < n_q; ++c) {
        if (used[c]) continue;
        if (pr->ns >= PM_MAX_SITES) { delete pr; return fail(PM_ERR_ARG, "pm_probe_create: too many sites"); }
        Site s{};
        s.ox = q_xyz[3 * c]; s.oy = q_xyz[3 * c + 1]; s.oz = q_xyz[3 * c + 2];
        s.lj = 0;
        s.q = q[c];
        pr->lj_species[pr->ns] = 0;
        pr->site[pr->ns++] = s;
    }
    for (int i = 0; i < pr->ns; ++i) {
        const Site& s = pr->site[i];
        if (!std::isfinite(s.ox) || !std::isfinite(s.oy) || !std::isfinite(s.oz) || !std::isfinite(s.q)) {
            delete pr;
            return fail(PM_ERR_ARG, "pm_probe_create: non-finite site");
        }
        if (s.q != 0.0) pr->charged = true;
    }
    if (pr->ns == 7) {  // no 7-site instantiation: pad with an inert site
        Site s{};
        pr->lj_species[7] = 0;
        pr->site[7] = s;
        pr->ns = 8;
    }
    *out = pr;
    return PM_OK;
}

int pm_probe_destroy(pm_probe* pr) {
    delete pr;
    return PM_OK;
}

}  // extern "C"

namespace {

// Shared set-up for the three modes: box and cut-offs into the kernel
// arguments, the LJ table into the head of the scratch buffer, and a staging
// region of `extra_bytes` right after it (returned in *extra) for the
// caller's inputs and outputs.
int prepare_energy(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, KArgs& a,
                   const int32_t* site_species, int precision, bool es, size_t extra_bytes,
                   unsigned char** extra) {
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    fill_box(a, fw);
    // no copy from host_tab may still be in flight (a failed call can return
    // before its final synchronisation)
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    std::vector<unsigned char>& tab = g_ctx.host_tab;
    if (precision == PM_FP64) build_ljtab<double>(fw, ff, a.ns, a.site, site_species, tab);
    else build_ljtab<float>(fw, ff, a.ns, a.site, site_species, tab);
    size_t tab_bytes = (tab.size() + 255) & ~(size_t)255;
    if (int rc = ensure_scratch(tab_bytes + extra_bytes)) return rc;
    unsigned char* base = static_cast<unsigned char*>(g_ctx.scratch);
    PM_HIP(hipMemcpyAsync(base, tab.data(), tab.size(), hipMemcpyHostToDevice, g_ctx.stream));
    a.ljtab = base;
    *extra = base + tab_bytes;
    a.r2_lj = ff ? ff->r2_cutoff : 0.0;
    if (es) {
        a.alpha = ew->alpha;
        a.r2_sr = ew->sr_cutoff * ew->sr_cutoff;
        a.kv = ew->d_k;
        a.nk = (int)ew->nk;
    } else {
        a.alpha = 1.0;
        a.r2_sr = 0.0;
        a.kv = nullptr;
        a.nk = 0;
    }
    return PM_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Explicit-pose driver behind pm_vdw_energy and pm_electrostatic_energy:
// out3 receives (vdw, sr, lr) per pose.
int run_poses(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, KArgs& a,
              const int32_t* site_species, int precision, bool es, const double* site_xf,
              std::vector<double>& out3) {
    int64_t n = a.n_items;
    size_t pose_bytes = align256((size_t)n * a.ns * 3 * sizeof(double));
    size_t out_bytes = align256((size_t)n * 3 * sizeof(double));
    unsigned char* extra = nullptr;
    if (int rc = prepare_energy(fw, ff, ew, a, site_species, precision, es, pose_bytes + out_bytes, &extra))
        return rc;
    double* d_pose = reinterpret_cast<double*>(extra);
    double* d_out = reinterpret_cast<double*>(extra + pose_bytes);
    PM_HIP(hipMemcpyAsync(d_pose, site_xf, (size_t)n * a.ns * 3 * sizeof(double), hipMemcpyHostToDevice,
                          g_ctx.stream));
    a.pose_xf = d_pose;
    a.out_e = d_out;
    int n_wg = (int)((n + WG - 1) / WG);
    if (int rc = launch_energy(a, es, precision, n_wg)) return rc;
    out3.resize((size_t)n * 3);
    PM_HIP(hipMemcpyAsync(out3.data(), d_out, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                          g_ctx.stream));
    return finish_timing((double)n * a.ns * fw->n_rep);
}

}  // namespace

extern "C" {

int pm_vdw_energy(const pm_framework* fw, const pm_forcefield* ff, int32_t n_sites,
                  const int32_t* site_species, int64_t n_poses, const double* site_xf,
                  int32_t precision, double* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !ff || !out || n_poses < 0 || n_sites < 1 || n_sites > PM_MAX_SITES || n_sites == 7 ||
        !site_species || (n_poses > 0 && !site_xf))
        return fail(PM_ERR_ARG, "pm_vdw_energy: bad handle, pointer or size (sites in 1..6 or 8)");
    if (int rc = check_species(fw, ff)) return rc;
    for (int s = 0; s < n_sites; ++s)
        if (site_species[s] < 0 || site_species[s] >= ff->n)
            return fail(PM_ERR_ARG, "pm_vdw_energy: site species outside the force field");
    if (n_poses == 0) return PM_OK;
    if (n_poses > (int64_t)INT32_MAX - WG) return fail(PM_ERR_ARG, "pm_vdw_energy: too many poses");
    KArgs a{};
    a.mode = MODE_POSES;
    a.ns = n_sites;
    a.n_items = n_poses;
    for (int s = 0; s < n_sites; ++s) { a.site[s].lj = 1; a.site[s].q = 0.0; }
    std::vector<double> out3;
    if (int rc = run_poses(fw, ff, nullptr, a, site_species, precision, false, site_xf, out3)) return rc;
    for (int64_t i = 0; i < n_poses; ++i) out[i] = out3[3 * i];
    return PM_OK;
}

int pm_electrostatic_energy(const pm_framework* fw, const pm_ewald* ew, int32_t n_charges,
                            const double* charges, int64_t n_poses, const double* charge_xf,
                            int32_t precision, double* out_sr, double* out_lr) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !ew || !out_sr || !out_lr || n_poses < 0 || n_charges < 1 ||
        n_charges > PM_MAX_SITES || n_charges == 7 || !charges || (n_poses > 0 && !charge_xf))
        return fail(PM_ERR_ARG, "pm_electrostatic_energy: bad handle, pointer or size (charges in 1..6 or 8)");
    for (int s = 0; s < n_charges; ++s)
        if (!std::isfinite(charges[s])) return fail(PM_ERR_ARG, "pm_electrostatic_energy: non-finite charge");
    if (n_poses == 0) return PM_OK;
    if (n_poses > (int64_t)INT32_MAX - WG) return fail(PM_ERR_ARG, "pm_electrostatic_energy: too many poses");
    KArgs a{};
    a.mode = MODE_POSES;
    a.ns = n_charges;
    a.n_items = n_poses;
    std::vector<int32_t> zero_species(n_charges, 0);
    for (int s = 0; s < n_charges; ++s) { a.site[s].lj = 0; a.site[s].q = charges[s]; }
    bool es = fw->charged;
    std::vector<double> out3;
    if (int rc = run_poses(fw, nullptr, es ? ew : nullptr, a, zero_species.data(), precision, es, charge_xf, out3))
        return rc;
    for (int64_t i = 0; i < n_poses; ++i) {
        out_sr[i] = out3[3 * i + 1];
        out_lr[i] = out3[3 * i + 2];
    }
    return PM_OK;
}

int pm_widom(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
             const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
             int64_t n_per_block, int32_t precision, double* out_block_sums,
             double* out_energies) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !ff || !pr || !out_block_sums || n_blocks < 1 || n_per_block < 1)
        return fail(PM_ERR_ARG, "pm_widom: null handle/pointer or empty block layout");
    if (!(temperature > 0.0) || !std::isfinite(temperature))
        return fail(PM_ERR_ARG, "pm_widom: temperature must be positive and finite");
    if (int rc = check_species(fw, ff)) return rc;
    for (int s = 0; s < pr->ns; ++s)
        if (pr->site[s].lj && pr->lj_species[s] >= ff->n)
            return fail(PM_ERR_ARG, "pm_widom: probe species outside the force field");
    bool es = fw->charged && pr->charged;  // src/henry.jl:93
    if (es && !ew) return fail(PM_ERR_ARG, "pm_widom: charged framework and probe need an Ewald handle");
    int64_t chunks = (n_per_block + WG - 1) / WG;
    int64_t n_wg = chunks * n_blocks;
    if (n_wg > (int64_t)INT32_MAX || (int64_t)n_blocks * n_per_block > ((int64_t)1 << 53))
        return fail(PM_ERR_ARG, "pm_widom: too many insertions for one call");
    KArgs a{};
    a.mode = MODE_WIDOM;
    a.ns = pr->ns;
    a.rotate = pr->rotate;
    a.n_blocks = n_blocks;
    a.chunks_per_block = (int)chunks;
    a.n_per_block = n_per_block;
    a.n_items = (long long)n_blocks * n_per_block;
    a.temperature = temperature;
    a.seed = seed;
    for (int s = 0; s < pr->ns; ++s) a.site[s] = pr->site[s];
    size_t part_bytes = align256((size_t)n_wg * 3 * sizeof(double));
    size_t sums_bytes = align256((size_t)n_blocks * 3 * sizeof(double));
    double* d_energies = nullptr;
    if (out_energies) {
        hipError_t e = hipMalloc(&d_energies, (size_t)a.n_items * 2 * sizeof(double));
        if (e != hipSuccess) return fail(PM_ERR_OOM, "pm_widom: energy buffer allocation failed");
    }
    a.out_e = d_energies;
    unsigned char* extra = nullptr;
    int rc = prepare_energy(fw, ff, es ? ew : nullptr, a, pr->lj_species, precision, es,
                            part_bytes + sums_bytes, &extra);
    if (rc == PM_OK) {
        a.partials = reinterpret_cast<double*>(extra);
        rc = launch_energy(a, es, precision, (int)n_wg);
    }
    if (rc != PM_OK) { hipFree(d_energies); return rc; }
    double* d_part = reinterpret_cast<double*>(extra);
    double* d_sums = reinterpret_cast<double*>(extra + part_bytes);
    hipLaunchKernelGGL(block_sums_kernel, dim3((n_blocks + 63) / 64), dim3(64), 0, g_ctx.stream,
                       d_part, (int)chunks, n_blocks, d_sums);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) { hipFree(d_energies); return fail(PM_ERR_DEVICE, hipGetErrorString(le)); }
    hipError_t ce = hipMemcpyAsync(out_block_sums, d_sums, (size_t)n_blocks * 3 * sizeof(double),
                                   hipMemcpyDeviceToHost, g_ctx.stream);
    if (ce == hipSuccess && out_energies)
        ce = hipMemcpyAsync(out_energies, d_energies, (size_t)a.n_items * 2 * sizeof(double),
                            hipMemcpyDeviceToHost, g_ctx.stream);
    int frc = finish_timing((double)a.n_items * pr->ns * fw->n_rep);
    hipFree(d_energies);
    if (ce != hipSuccess) return fail(PM_ERR_DEVICE, std::string("pm_widom: copy: ") + hipGetErrorString(ce));
    return frc;
}

int pm_energy_grid(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                   const pm_probe* pr, const int32_t n_pts[3], int32_t precision, double* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !pr || !n_pts || !out)
        return fail(PM_ERR_ARG, "pm_energy_grid: null handle or pointer");
    if (pr->ns != 1) return fail(PM_ERR_ARG, "pm_energy_grid: probe must have exactly one site");
    if (pr->site[0].ox != 0.0 || pr->site[0].oy != 0.0 || pr->site[0].oz != 0.0)
        return fail(PM_ERR_ARG, "pm_energy_grid: single-site probe must sit at its centre of mass");
    for (int i = 0; i < 3; ++i)
        if (n_pts[i] < 1 || n_pts[i] > 65536) return fail(PM_ERR_ARG, "pm_energy_grid: n_pts out of range");
    if (pr->site[0].lj) {
        if (!ff) return fail(PM_ERR_ARG, "pm_energy_grid: LJ probe needs a force field");
        if (int rc = check_species(fw, ff)) return rc;
        if (pr->lj_species[0] >= ff->n) return fail(PM_ERR_ARG, "pm_energy_grid: probe species outside the force field");
    }
    bool es = fw->charged && pr->charged;
    if (es && !ew) return fail(PM_ERR_ARG, "pm_energy_grid: charged probe needs an Ewald handle");
    long long n_items = (long long)n_pts[0] * n_pts[1] * n_pts[2];
    if ((n_items + WG - 1) / WG > INT32_MAX) return fail(PM_ERR_ARG, "pm_energy_grid: grid too large");
    KArgs a{};
    a.mode = MODE_GRID;
    a.ns = 1;
    a.n_items = n_items;
    for (int i = 0; i < 3; ++i) a.grid_n[i] = n_pts[i];
    a.site[0] = pr->site[0];
    bool fp64 = precision == PM_FP64;
    void* d_grid = nullptr;
    hipError_t e = hipMalloc(&d_grid, (size_t)n_items * (fp64 ? sizeof(double) : sizeof(float)));
    if (e != hipSuccess) return fail(PM_ERR_OOM, "pm_energy_grid: grid allocation failed");
    if (fp64) a.out_e = static_cast<double*>(d_grid);
    else a.out_f = static_cast<float*>(d_grid);
    int n_wg = (int)((n_items + WG - 1) / WG);
    unsigned char* extra = nullptr;
    int rc = prepare_energy(fw, ff, es ? ew : nullptr, a, pr->lj_species, precision, es, 0, &extra);
    if (rc == PM_OK) rc = launch_energy(a, es, precision, n_wg);
    if (rc != PM_OK) { hipFree(d_grid); return rc; }
    hipError_t ce;
    if (fp64) {
        ce = hipMemcpyAsync(out, d_grid, (size_t)n_items * sizeof(double), hipMemcpyDeviceToHost, g_ctx.stream);
        int frc = finish_timing((double)n_items * fw->n_rep);
        hipFree(d_grid);
        if (ce != hipSuccess) return fail(PM_ERR_DEVICE, hipGetErrorString(ce));
        return frc;
    }
    std::vector<float> tmp((size_t)n_items);
    ce = hipMemcpyAsync(tmp.data(), d_grid, (size_t)n_items * sizeof(float), hipMemcpyDeviceToHost, g_ctx.stream);
    int frc = finish_timing((double)n_items * fw->n_rep);
    hipFree(d_grid);
    if (ce != hipSuccess) return fail(PM_ERR_DEVICE, hipGetErrorString(ce));
    for (long long i = 0; i < n_items; ++i) out[i] = (double)tmp[i];
    return frc;
}

int pm_last_kernel_stats(double* kernel_ms, double* pair_evals) {
    if (kernel_ms) *kernel_ms = g_ctx.last_ms;
    if (pair_evals) *pair_evals = g_ctx.last_pairs;
    return PM_OK;
}

}  // extern "C"
