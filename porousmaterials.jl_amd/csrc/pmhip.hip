// pmhip.hip -- host side and C ABI of libpmhip.so (declared in include/pmhip.h).
//
// Device code is in pm_device.h.  Handles own device copies of their data;
// every entry point validates its arguments on the host, launches on the
// library stream and synchronises before returning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "pm_gcmc.h"
#include "pm_pockets.h"
#include "pm_points.h"
#include "pmhip.h"

#define PM_VERSION "pmhip 0.2.0 (gfx950)"

using namespace pmk;

// Scratch of pm_widom_noalloc, the entry point that allocates nothing: every
// GPU that loads the library holds this array from module load until the
// process ends.  Host code never uses the name itself as an address (on the
// host it is a stand-in, not device memory): pm_init resolves its device
// address once, with hipGetSymbolAddress, and everything uses that.
static __device__ __align__(256) unsigned char scratch[PM_NOALLOC_SCRATCH_BYTES];

namespace {

// Point-engine tunables.
// Point-engine tunables (defaults; pm_set_tuning / pm_autotune_widom change them).
constexpr double DEFAULT_CHUNK_R_MAX = 6.0;   // A: largest chunk radius of the image-list path
constexpr double DEFAULT_CELL_POINTS = 320.0; // points in the cube that sets a Widom cell's cross-section
constexpr long long MAX_POINT_SLOTS = 1LL << 24;  // point slots per launch group
constexpr long long MAX_IMAGES = 1LL << 26;       // images per list
constexpr size_t MAX_IMAGE_LISTS = 8;             // image lists (radii) per framework
constexpr long long MAX_GROUP_CELLS = 1LL << 24;  // cells of one launch group
constexpr double SB_POINTS = 32768.0;             // target points per super-bin of the Widom sort
constexpr double ZSLAB_POINTS = 64.0;             // target points per slab of a Widom cell
constexpr int MAX_RUN_CELLS = 65536;              // cells of one Widom run
static_assert(MAX_POINT_SLOTS <= (long long)pmk::SB_SLOT_MASK + 1, "a super-bin slot fits below the copy bits");
static_assert(MAX_GROUP_CELLS <= (long long)pmk::CELL_MASK + 1, "a group's cell index fits below the slab bits");
static_assert(pmk::ZBINS_MAX < 256 && pmk::SB_KEYS_MAX / 16 >= pmk::ZBINS_MAX,
              "slabs fit the top byte of the cell index, and 16 cells of them a super-bin");
static_assert(MAX_RUN_CELLS / 16 + 1 <= pmk::SB_SPAN_MAX, "a run's cells span few enough super-bins of 16 cells");

thread_local std::string g_err;
std::mutex g_mu;

struct Ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evp0 = nullptr, evp1 = nullptr;
    double last_ms = 0.0, last_pair_ms = 0.0, last_pairs = 0.0, last_cand = 0.0;
    size_t max_lds = 0;  // LDS bytes one workgroup may allocate
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    void* noalloc_scratch = nullptr;  // device address of ::scratch on this device
    // pinned host staging of grid downloads and the events that mark its pieces
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    std::vector<hipEvent_t> piece_ev;
    // host staging for the per-call LJ table: outlives the asynchronous copy
    std::vector<unsigned char> host_tab;
    // pinned host staging of the point engine's per-call upload (run records,
    // segments, tables) and download (candidate count, block sums)
    void* io_pin = nullptr;
    size_t io_pin_bytes = 0;
    double cell_points = DEFAULT_CELL_POINTS;
    double chunk_r_max = DEFAULT_CHUNK_R_MAX;
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
    size_t want = std::max(bytes, (size_t)1 << 20);
    PM_HIP(hipMalloc(&g_ctx.scratch, want));
    g_ctx.scratch_bytes = want;
    return PM_OK;
}

// Pinned host staging for grid downloads (grown on demand, held until the
// process ends).
int ensure_pinned(size_t bytes) {
    if (bytes <= g_ctx.pinned_bytes) return PM_OK;
    if (g_ctx.pinned) {
        PM_HIP(hipStreamSynchronize(g_ctx.stream));
        PM_HIP(hipHostFree(g_ctx.pinned));
        g_ctx.pinned = nullptr;
        g_ctx.pinned_bytes = 0;
    }
    const size_t want = std::max(bytes, (size_t)1 << 24);
    PM_HIP(hipHostMalloc(&g_ctx.pinned, want, hipHostMallocDefault));
    g_ctx.pinned_bytes = want;
    return PM_OK;
}

// The pinned upload/download staging of the point engine, grown on demand when
// `may_alloc` (held until the process ends); nullptr when it is smaller than
// `bytes` and may not grow, and the caller then copies through pageable memory.
int io_staging(size_t bytes, bool may_alloc, unsigned char** out) {
    *out = nullptr;
    if (bytes > g_ctx.io_pin_bytes && may_alloc) {
        if (g_ctx.io_pin) {
            PM_HIP(hipStreamSynchronize(g_ctx.stream));
            PM_HIP(hipHostFree(g_ctx.io_pin));
            g_ctx.io_pin = nullptr;
            g_ctx.io_pin_bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)1 << 20);
        PM_HIP(hipHostMalloc(&g_ctx.io_pin, want, hipHostMallocDefault));
        g_ctx.io_pin_bytes = want;
    }
    if (bytes <= g_ctx.io_pin_bytes) *out = static_cast<unsigned char*>(g_ctx.io_pin);
    return PM_OK;
}

int ensure_piece_events(size_t n) {
    while (g_ctx.piece_ev.size() < n) {
        hipEvent_t e = nullptr;
        PM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        g_ctx.piece_ev.push_back(e);
    }
    return PM_OK;
}

// One piece of a grid download: elements [e0, e0 + n) of a grid, from the
// device grid (FP32 or FP64) through the pinned staging into the caller's
// FP64 buffer.
struct GridPiece {
    const unsigned char* dev;
    unsigned char* pin;
    double* dst;
    long long n;
};

constexpr size_t PIECE_BYTES = (size_t)4 << 20;   // bytes of one download piece
constexpr size_t MAX_PINNED = (size_t)2 << 30;    // larger downloads copy to the caller directly
constexpr unsigned MAX_HOST_THREADS = 16;

// Enqueues the downloads of every piece (each followed by its event) and
// widens / copies them into the callers' buffers on up to MAX_HOST_THREADS
// host threads as the pieces arrive, so the copy over the host link overlaps
// the host-side conversion.  Returns once every piece is in place.
int download_grids(std::vector<GridPiece>& pieces, bool f32) {
    if (pieces.empty()) return PM_OK;
    if (int rc = ensure_piece_events(pieces.size())) return rc;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    for (size_t k = 0; k < pieces.size(); ++k) {
        PM_HIP(hipMemcpyAsync(pieces[k].pin, pieces[k].dev, (size_t)pieces[k].n * es, hipMemcpyDeviceToHost,
                              g_ctx.stream));
        PM_HIP(hipEventRecord(g_ctx.piece_ev[k], g_ctx.stream));
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<size_t>(std::min(hw, MAX_HOST_THREADS), pieces.size());
    std::vector<hipError_t> errs(nt, hipSuccess);
    auto work = [&](unsigned t) {
        for (size_t k = t; k < pieces.size(); k += nt) {
            const hipError_t e = hipEventSynchronize(g_ctx.piece_ev[k]);
            if (e != hipSuccess) {
                errs[t] = e;
                return;
            }
            const GridPiece& pc = pieces[k];
            if (f32) {
                const float* src = reinterpret_cast<const float*>(pc.pin);
                for (long long i = 0; i < pc.n; ++i) pc.dst[i] = (double)src[i];
            } else {
                std::memcpy(pc.dst, pc.pin, (size_t)pc.n * sizeof(double));
            }
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(nt);
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (hipError_t e : errs)
        if (e != hipSuccess) return fail(PM_ERR_DEVICE, std::string("grid download: ") + hipGetErrorString(e));
    return PM_OK;
}

int require_init() {
    if (g_ctx.device < 0) return fail(PM_ERR_STATE, "pm_init has not been called");
    return PM_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Carves aligned sub-buffers out of one scratch allocation.
struct Arena {
    size_t used = 0;
    size_t take(size_t bytes) {
        size_t off = used;
        used += align256(bytes);
        return off;
    }
};

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

void cross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

double norm3(const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Perpendicular widths of the parallelepiped spanned by the columns of F.
void widths(const double F[9], double w[3]) {
    double n[3];
    double vol = std::fabs(F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
                           F[6] * (F[1] * F[5] - F[4] * F[2]));
    cross(&F[3], &F[6], n);
    w[0] = vol / norm3(n);
    cross(&F[0], &F[6], n);
    w[1] = vol / norm3(n);
    cross(&F[0], &F[3], n);
    w[2] = vol / norm3(n);
}

// Half of the longest diagonal of the parallelepiped with edges ea, eb, ec.
double half_diagonal(const double* ea, const double* eb, const double* ec) {
    double best = 0.0;
    for (int sb = -1; sb <= 1; sb += 2)
        for (int sc = -1; sc <= 1; sc += 2) {
            double d[3];
            for (int i = 0; i < 3; ++i) d[i] = ea[i] + sb * eb[i] + sc * ec[i];
            best = std::max(best, norm3(d));
        }
    return 0.5 * best;
}

const double INV_4PI_EPS0 =
    1.0 / (4.0 * M_PI * (8.854187817e-12 / (1.602176565e-19 * 1.602176565e-19) / 1.0e10 * 1.3806488e-23));

}  // namespace

struct pm_framework {
    double F[9];      // replicated-box f_to_c
    double c2f[9];
    int32_t rep[3];
    double width[3];  // perpendicular widths of the replicated box
    int64_t n_unit = 0, n_rep = 0, n_unit_q = 0;
    std::vector<int32_t> species_global;  // compact -> force-field species
    std::vector<double> unit_xf, unit_q;  // host copies for the image list
    std::vector<int32_t> unit_type;
    double net_charge = 0.0;
    bool charged = false;
    double4* d_atoms = nullptr;
    int32_t* d_species = nullptr;
    double4* d_unit_q = nullptr;
    // Image lists, one per cutoff radius r: every periodic image within r of
    // some point of the replicated box, Cartesian xyz + charge, and its type.
    struct Images {
        double r = 0.0;
        int64_t n = 0;
        int64_t n_tiles = 0;
        double4* d_img = nullptr;
        int32_t* d_type = nullptr;
        double4* d_tile_sph = nullptr;  // tile bounding spheres
        int2* d_tile_rng = nullptr;     // tile (first image, count)
    };
    std::vector<Images> images;
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
    int Ma = 0, Mb = 0, Mc = 0;
    double2* d_box = nullptr;
    int2* d_rows = nullptr;    // per (ma, mb) row of the box: kept mc range (lo > hi: empty)
    float4* d_erfc = nullptr;  // ERFC_SEG cubics of erfc(alpha r) on [0, sr_cutoff]
    double2* d_erfc64 = nullptr;  // ERFC64_SEG quintics of erfc(alpha r) on [0, sr_cutoff] (FP64)
    double2* d_box_t = nullptr;  // per-type boxes (gradient runs), built on first use
};

struct pm_probe {
    int ns = 0;
    int rotate = 0;
    Site site[PM_MAX_SITES];
    int32_t lj_species[PM_MAX_SITES];
    bool charged = false;
};

struct pm_access {
    int32_t n[3] = {0, 0, 0};
    unsigned char* d = nullptr;  // 0 / 1 per unit-cell grid point, i fastest
};

struct pm_gcmc {
    int n_chains = 0, n_species = 0, capacity = 0;
    int64_t n_dens = 0;           // density-grid points per chain (0: no grid)
    bool energy_stale = false;    // a starting configuration changed the system energy
    GcmcParams hp{};              // host copy of the shared parameters
    GcmcParams* d_params = nullptr;
    GcmcChain* d_chains = nullptr;
    double4* d_atoms = nullptr;
    double2* d_fg = nullptr;
    double2* d_gg = nullptr;
    double4* d_sites = nullptr;
    double4* d_com = nullptr;
    unsigned* d_dens = nullptr;
    double* d_tot = nullptr;
    int* d_incomplete = nullptr;  // chains a launch with molecules in LDS left unfinished
    int* d_cell_start = nullptr;  // framework cell lists (GcmcParams::cell_start, cell_atoms)
    int* d_cell_atoms = nullptr;
    bool mols_in_memory = false;  // a chain outgrew the LDS molecule store: launches keep them in memory
    int lds_mols_limit = -1;      // pm_gcmc_set_lds_molecules
};

namespace {

// LJ table [n_types][ns] of (sigma^2, 4 eps) for the framework's compact
// species against each probe site, in the precision of the pair maths.
template <typename Real>
void build_ljtab(const pm_framework* fw, const pm_forcefield* ff, int ns, const Site* site,
                 const int32_t* site_species, std::vector<unsigned char>& out) {
    int nt = (int)fw->species_global.size();
    out.assign(sizeof(LJP<Real>) * (size_t)(nt > 0 ? nt : 1) * ns, 0);
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

// Builds the LJ table in the host staging buffer; returns its byte size.
int stage_ljtab(const pm_framework* fw, const pm_forcefield* ff, int ns, const Site* site,
                const int32_t* site_species, int precision, size_t* bytes) {
    // no copy from host_tab may still be in flight (a failed call can return
    // before its final synchronisation)
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    std::vector<unsigned char>& tab = g_ctx.host_tab;
    if (precision == PM_FP64) build_ljtab<double>(fw, ff, ns, site, site_species, tab);
    else build_ljtab<float>(fw, ff, ns, site, site_species, tab);
    *bytes = tab.size();
    return PM_OK;
}

int copy_ljtab(void* dst) {
    PM_HIP(hipMemcpyAsync(dst, g_ctx.host_tab.data(), g_ctx.host_tab.size(), hipMemcpyHostToDevice,
                          g_ctx.stream));
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

// ----------------------------------------------------------- poses ----
size_t pose_lds_bytes(int n_types, int ns, bool fp64) {
    size_t region = std::max((size_t)TILE * sizeof(double4), (size_t)KTILE * sizeof(KVec));
    return region + TILE * sizeof(int) + (size_t)std::max(n_types, 1) * ns * (fp64 ? 16 : 8);
}

template <int NS, bool ES, typename Real, bool GRAD>
void pose_launch_t(const PoseArgs& a, int n_wg, size_t shm) {
    hipLaunchKernelGGL((pose_kernel<NS, ES, Real, GRAD>), dim3(n_wg), dim3(WG), shm, g_ctx.stream, a);
}

template <bool ES, typename Real, bool GRAD>
int pose_launch_ns(const PoseArgs& a, int n_wg, size_t shm) {
    switch (a.ns) {
        case 1: pose_launch_t<1, ES, Real, GRAD>(a, n_wg, shm); break;
        case 2: pose_launch_t<2, ES, Real, GRAD>(a, n_wg, shm); break;
        case 3: pose_launch_t<3, ES, Real, GRAD>(a, n_wg, shm); break;
        case 4: pose_launch_t<4, ES, Real, GRAD>(a, n_wg, shm); break;
        case 5: pose_launch_t<5, ES, Real, GRAD>(a, n_wg, shm); break;
        case 6: pose_launch_t<6, ES, Real, GRAD>(a, n_wg, shm); break;
        case 8: pose_launch_t<8, ES, Real, GRAD>(a, n_wg, shm); break;
        default: return fail(PM_ERR_ARG, "unsupported site count " + std::to_string(a.ns));
    }
    return PM_OK;
}

template <bool GRAD>
int pose_launch(const PoseArgs& a, int n_wg, size_t shm, bool fp64, bool es) {
    if (fp64) return es ? pose_launch_ns<true, double, GRAD>(a, n_wg, shm) : pose_launch_ns<false, double, GRAD>(a, n_wg, shm);
    return es ? pose_launch_ns<true, float, GRAD>(a, n_wg, shm) : pose_launch_ns<false, float, GRAD>(a, n_wg, shm);
}

// Explicit-pose driver behind pm_vdw_energy, pm_electrostatic_energy and
// pm_pose_gradient: out3 receives (vdw, sr, lr) per pose; out_grad (if not
// null) n_poses x ns x 3 site gradients.
int run_poses(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, PoseArgs a,
              const int32_t* site_species, int precision, bool es, const double* site_xf,
              std::vector<double>& out3, double* out_grad = nullptr) {
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    const int64_t n = a.n_items;
    a.f11 = fw->F[0]; a.f12 = fw->F[3]; a.f13 = fw->F[6];
    a.f22 = fw->F[4]; a.f23 = fw->F[7]; a.f33 = fw->F[8];
    for (int i = 0; i < 3; ++i) a.repd[i] = (double)fw->rep[i];
    for (int i = 0; i < 9; ++i) a.c2f[i] = fw->c2f[i];
    a.n_atoms = (int)fw->n_rep;
    a.n_types = (int)fw->species_global.size();
    a.atoms = fw->d_atoms;
    a.species = fw->d_species;
    a.inv4pe = INV_4PI_EPS0;
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
    size_t tab_bytes = 0;
    if (int rc = stage_ljtab(fw, ff, a.ns, a.site, site_species, precision, &tab_bytes)) return rc;
    Arena ar;
    size_t o_tab = ar.take(tab_bytes);
    size_t o_pose = ar.take((size_t)n * a.ns * 3 * sizeof(double));
    size_t o_out = ar.take((size_t)n * 3 * sizeof(double));
    size_t o_grad = ar.take(out_grad ? (size_t)n * a.ns * 3 * sizeof(double) : 0);
    if (int rc = ensure_scratch(ar.used)) return rc;
    unsigned char* base = static_cast<unsigned char*>(g_ctx.scratch);
    if (int rc = copy_ljtab(base + o_tab)) return rc;
    a.ljtab = base + o_tab;
    double* d_pose = reinterpret_cast<double*>(base + o_pose);
    double* d_out = reinterpret_cast<double*>(base + o_out);
    PM_HIP(hipMemcpyAsync(d_pose, site_xf, (size_t)n * a.ns * 3 * sizeof(double), hipMemcpyHostToDevice,
                          g_ctx.stream));
    a.pose_xf = d_pose;
    a.out_e = d_out;
    a.out_g = out_grad ? reinterpret_cast<double*>(base + o_grad) : nullptr;
    const bool fp64 = precision == PM_FP64;
    const size_t shm = pose_lds_bytes(a.n_types, a.ns, fp64);
    const int n_wg = (int)((n + WG / 64 - 1) / (WG / 64));  // one wave per pose
    PM_HIP(hipEventRecord(g_ctx.ev0, g_ctx.stream));
    const int rc = out_grad ? pose_launch<true>(a, n_wg, shm, fp64, es) : pose_launch<false>(a, n_wg, shm, fp64, es);
    if (rc != PM_OK) return rc;
    PM_HIP(hipGetLastError());
    PM_HIP(hipEventRecord(g_ctx.ev1, g_ctx.stream));
    out3.resize((size_t)n * 3);
    PM_HIP(hipMemcpyAsync(out3.data(), d_out, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                          g_ctx.stream));
    if (out_grad)
        PM_HIP(hipMemcpyAsync(out_grad, a.out_g, (size_t)n * a.ns * 3 * sizeof(double), hipMemcpyDeviceToHost,
                              g_ctx.stream));
    return finish_timing((double)n * a.ns * fw->n_rep);
}

// ---------------------------------------------------- point engine ----
// Image list for radius r: every periodic image (x_unit + t) / rep whose
// replicated-box fractional coordinates lie within [-m_a, 1 / rep_a + m_a] on
// each axis, m_a = r / width_a, so that every image within r of any point of
// the first unit cell [0, 1 / rep]^3 of the replicated box is present (the
// point engine's points all lie there: Widom sites are folded into it, grids
// span it); stored in Cartesian coordinates.  Lists are cached
// per framework and radius, so a run's list -- and with it the order of its
// FP32 partial sums -- depends on the framework and the radius alone.
const pm_framework::Images* find_images(const pm_framework* fw, double r) {
    for (const auto& im : fw->images)
        if (im.r == r) return &im;
    return nullptr;
}

int get_images(pm_framework* fw, double r, const pm_framework::Images** out) {
    if (const pm_framework::Images* im = find_images(fw, r)) {
        *out = im;
        return PM_OK;
    }
    if (fw->images.size() >= MAX_IMAGE_LISTS) return fail(PM_ERR_ARG, "too many image-list radii");
    double m[3];
    for (int i = 0; i < 3; ++i) m[i] = r / fw->width[i];
    int64_t est = fw->n_unit;
    for (int i = 0; i < 3; ++i) est *= (int64_t)std::ceil(1.0 + 2.0 * m[i] * fw->rep[i]) + 1;
    if (est > MAX_IMAGES) return fail(PM_ERR_ARG, "image list too large");
    std::vector<double4> img;
    std::vector<int32_t> ty;
    img.reserve((size_t)est);
    ty.reserve((size_t)est);
    const double* F = fw->F;
    // by species (then unit-cell atom) before the spatial ordering below, so
    // the species of a tile come in order and a gradient run's per-type sums
    // change type at most once per species in a tile
    std::vector<int64_t> by_type((size_t)fw->n_unit);
    for (int64_t u = 0; u < fw->n_unit; ++u) by_type[(size_t)u] = u;
    std::stable_sort(by_type.begin(), by_type.end(),
                     [&](int64_t x, int64_t y) { return fw->unit_type[x] < fw->unit_type[y]; });
    for (int64_t u : by_type) {
        const double* x = &fw->unit_xf[3 * u];
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = (int)std::ceil(-m[a] * fw->rep[a] - x[a]);
            hi[a] = (int)std::floor(1.0 + m[a] * fw->rep[a] - x[a]);
        }
        for (int ta = lo[0]; ta <= hi[0]; ++ta)
            for (int tb = lo[1]; tb <= hi[1]; ++tb)
                for (int tc = lo[2]; tc <= hi[2]; ++tc) {
                    const double fa = (x[0] + ta) / fw->rep[0], fb = (x[1] + tb) / fw->rep[1],
                                 fc = (x[2] + tc) / fw->rep[2];
                    img.push_back(make_double4(F[0] * fa + F[3] * fb + F[6] * fc, F[4] * fb + F[7] * fc,
                                               F[8] * fc, fw->unit_q[u]));
                    ty.push_back(fw->unit_type[u]);
                }
    }
    // spatial tiles: bins of about TILE_IMAGES images over the list's bounding
    // box; the list is ordered by (bin, species, previous order) and every bin
    // is cut into tiles of at most TILE_IMAGES images, each with its bounding
    // sphere.  A chunk visits only the tiles whose sphere meets its own; the
    // FP32 partial sums of a point run per tile, so they depend on this list
    // alone.
    std::vector<double4> sph;
    std::vector<int2> rng;
    if (!img.empty()) {
        double lo[3] = {img[0].x, img[0].y, img[0].z}, hi[3] = {lo[0], lo[1], lo[2]};
        for (const auto& v : img) {
            const double c[3] = {v.x, v.y, v.z};
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], c[d]);
                hi[d] = std::max(hi[d], c[d]);
            }
        }
        double ext[3], bvol = 1.0;
        for (int d = 0; d < 3; ++d) {
            ext[d] = std::max(hi[d] - lo[d], 1e-6);
            bvol *= ext[d];
        }
        const double b = std::cbrt(bvol * TILE_IMAGES / (double)img.size());
        int nb[3];
        for (int d = 0; d < 3; ++d) nb[d] = (int)std::max(1.0, std::min(4096.0, std::ceil(ext[d] / b)));
        std::vector<int64_t> key(img.size());
        for (size_t i = 0; i < img.size(); ++i) {
            const double c[3] = {img[i].x, img[i].y, img[i].z};
            int64_t k = 0;
            for (int d = 0; d < 3; ++d) {
                const int q = std::min(nb[d] - 1, std::max(0, (int)((c[d] - lo[d]) / ext[d] * nb[d])));
                k = k * nb[d] + q;
            }
            key[i] = k;
        }
        std::vector<size_t> perm(img.size());
        for (size_t i = 0; i < perm.size(); ++i) perm[i] = i;
        std::stable_sort(perm.begin(), perm.end(), [&](size_t x, size_t y) {
            return key[x] != key[y] ? key[x] < key[y] : ty[x] < ty[y];
        });
        std::vector<double4> img2(img.size());
        std::vector<int32_t> ty2(ty.size());
        for (size_t i = 0; i < perm.size(); ++i) {
            img2[i] = img[perm[i]];
            ty2[i] = ty[perm[i]];
        }
        img.swap(img2);
        ty.swap(ty2);
        size_t i0 = 0;
        while (i0 < img.size()) {
            size_t i1 = i0 + 1;
            while (i1 < img.size() && i1 - i0 < (size_t)TILE_IMAGES && key[perm[i1]] == key[perm[i0]]) ++i1;
            double bl[3] = {img[i0].x, img[i0].y, img[i0].z}, bh[3] = {bl[0], bl[1], bl[2]};
            for (size_t i = i0; i < i1; ++i) {
                const double c[3] = {img[i].x, img[i].y, img[i].z};
                for (int d = 0; d < 3; ++d) {
                    bl[d] = std::min(bl[d], c[d]);
                    bh[d] = std::max(bh[d], c[d]);
                }
            }
            const double cx = 0.5 * (bl[0] + bh[0]), cy = 0.5 * (bl[1] + bh[1]), cz = 0.5 * (bl[2] + bh[2]);
            double r2 = 0.0;
            for (size_t i = i0; i < i1; ++i) {
                const double dx = img[i].x - cx, dy = img[i].y - cy, dz = img[i].z - cz;
                r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
            }
            // slack for the kernel's rounding of the sphere test
            sph.push_back(make_double4(cx, cy, cz, std::sqrt(r2) * (1.0 + 1e-12) + 1e-9));
            rng.push_back(make_int2((int)i0, (int)(i1 - i0)));
            i0 = i1;
        }
    }
    pm_framework::Images im;
    im.r = r;
    im.n = (int64_t)img.size();
    im.n_tiles = (int64_t)sph.size();
    if (im.n > (int64_t)INT32_MAX) return fail(PM_ERR_ARG, "image list too large");
    const size_t n = std::max<size_t>(img.size(), 1);
    const size_t nt = std::max<size_t>(sph.size(), 1);
    PM_HIP(hipMalloc(&im.d_img, n * sizeof(double4)));
    hipError_t e = hipMalloc(&im.d_type, n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&im.d_tile_sph, nt * sizeof(double4));
    if (e == hipSuccess) e = hipMalloc(&im.d_tile_rng, nt * sizeof(int2));
    if (e == hipSuccess && !img.empty()) {
        e = hipMemcpy(im.d_img, img.data(), img.size() * sizeof(double4), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(im.d_type, ty.data(), ty.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(im.d_tile_sph, sph.data(), sph.size() * sizeof(double4), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(im.d_tile_rng, rng.data(), rng.size() * sizeof(int2), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        (void)hipFree(im.d_img);
        (void)hipFree(im.d_type);
        (void)hipFree(im.d_tile_sph);
        (void)hipFree(im.d_tile_rng);
        return fail(PM_ERR_DEVICE, std::string("image list upload: ") + hipGetErrorString(e));
    }
    fw->images.push_back(im);
    *out = &fw->images.back();
    return PM_OK;
}

// Pair and Fourier parameters of a run.
void fill_run_pair(RunArgs& a, const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, bool es,
                   int nsite, const Site* site) {
    a.f11 = fw->F[0]; a.f12 = fw->F[3]; a.f13 = fw->F[6];
    a.f22 = fw->F[4]; a.f23 = fw->F[7]; a.f33 = fw->F[8];
    for (int i = 0; i < 9; ++i) a.c2f[i] = fw->c2f[i];
    for (int i = 0; i < 3; ++i) a.repd[i] = (double)fw->rep[i];
    a.r2_lj = ff ? ff->r2_cutoff : 0.0;
    a.r2_sr = es ? ew->sr_cutoff * ew->sr_cutoff : 0.0;
    a.alpha = es ? ew->alpha : 1.0;
    a.r_cmax = std::sqrt(std::max(a.r2_lj, a.r2_sr));
    a.erfc_tab = es ? ew->d_erfc : nullptr;
    a.erfc_inv_h = es ? ERFC_SEG / ew->sr_cutoff : 1.0;
    a.erfc64_tab = es ? ew->d_erfc64 : nullptr;
    a.erfc64_inv_h = es ? ERFC64_SEG / ew->sr_cutoff : 1.0;
    a.inv4pe = INV_4PI_EPS0;
    a.esm = !es ? 0 : (a.r2_lj == a.r2_sr ? 1 : 2);
    a.nsite = nsite;
    a.n_types = (int)fw->species_global.size();
    a.n_atoms = (int)fw->n_rep;
    a.atoms = fw->d_atoms;
    a.species = fw->d_species;
    a.site_lj_mask = 0;
    a.site_q_mask = 0;
    for (int s = 0; s < nsite; ++s) {
        if (site[s].lj) a.site_lj_mask |= 1u << s;
        if (site[s].q != 0.0) a.site_q_mask |= 1u << s;
    }
    if (es) {
        a.Ma = ew->Ma;
        a.Mb = ew->Mb;
        a.Mc = ew->Mc;
        a.A = ew->d_box;
        a.krows = ew->d_rows;
    }
}

// The cells of a Widom run over the first unit cell, from h, the edge of a
// cube holding cell_points points (the tuning knob).  Where such a cube holds
// at least a slice (WG points), the cells are columns along the third axis
// with an h x h cross-section and the full height of the unit cell (stacked
// only when a column would need more than ZBINS_MAX slabs), and within a cell
// the points are ordered by zbins slabs of about ZSLAB_POINTS points, so that
// WG consecutive points of a cell fill a compact piece of the column.  Sparser
// runs keep cubes of edge h and one slab.  chunk_r is the half diagonal of the
// piece of a cell that a slice fills (the scale of a workgroup's candidate
// sphere), which the image-list path needs at most chunk_r_max: h shrinks
// until it is, while the run has at most MAX_RUN_CELLS cells.
void choose_cells(const pm_framework* fw, double n_points, int ncell[3], double* chunk_r, int* zbins) {
    const double n_cells_box = (double)fw->rep[0] * fw->rep[1] * fw->rep[2];
    const double vol = std::fabs(fw->F[0] * fw->F[4] * fw->F[8]) / n_cells_box;
    const double n_pts = std::max(n_points, 1.0);
    double h = std::cbrt(vol * g_ctx.cell_points / n_pts);
    auto axis_cells = [&](int i, double edge) {
        return (int)std::min(1024.0, std::max(1.0, std::round(fw->width[i] / fw->rep[i] / edge)));
    };
    bool have = false;
    for (int iter = 0; iter < 64; ++iter) {
        int n[3] = {axis_cells(0, h), axis_cells(1, h), axis_cells(2, h)};
        int zb = 1;
        if (n_pts / ((double)n[0] * n[1] * n[2]) >= WG) {
            const double pts_col = n_pts / ((double)n[0] * n[1]);
            const long long zb_want = std::max(1LL, std::llround(pts_col / ZSLAB_POINTS));
            n[2] = (int)((zb_want + ZBINS_MAX - 1) / ZBINS_MAX);
            zb = (int)std::max(1LL, std::llround((double)zb_want / n[2]));
        }
        const double cells = (double)n[0] * n[1] * n[2];
        if (cells > (double)MAX_RUN_CELLS) {
            if (have) return;  // the finer cells of the previous step stand
            h *= 1.1;          // too fine from the start: coarser
            continue;
        }
        // the piece of a cell that WG consecutive points fill
        const double frac_c = std::min(1.0, (double)WG * cells / n_pts);
        double ea[3], eb[3], ec[3];
        for (int i = 0; i < 3; ++i) {
            ea[i] = fw->F[i] / (n[0] * fw->rep[0]);
            eb[i] = fw->F[3 + i] / (n[1] * fw->rep[1]);
            ec[i] = fw->F[6 + i] / (n[2] * fw->rep[2]) * frac_c;
        }
        for (int i = 0; i < 3; ++i) ncell[i] = n[i];
        *zbins = zb;
        *chunk_r = half_diagonal(ea, eb, ec);
        have = true;
        if (*chunk_r <= g_ctx.chunk_r_max) return;
        h /= 1.25;
    }
    if (!have) {  // (not reached for cell_points >= 1; one cell, all-atoms path)
        double ea[3], eb[3], ec[3];
        for (int i = 0; i < 3; ++i) {
            ea[i] = fw->F[i] / fw->rep[0];
            eb[i] = fw->F[3 + i] / fw->rep[1];
            ec[i] = fw->F[6 + i] / fw->rep[2];
        }
        ncell[0] = ncell[1] = ncell[2] = 1;
        *zbins = 1;
        *chunk_r = half_diagonal(ea, eb, ec);
    }
}

// One validated run of a point-engine call.
struct PointRun {
    pm_framework* fw = nullptr;
    const pm_forcefield* ff = nullptr;
    const pm_ewald* ew = nullptr;
    const pm_probe* pr = nullptr;
    bool es = false;
    int ns = 0;
    int64_t n_wg = 0;      // Widom: insertion workgroups; grid: bricks
    int64_t n_points = 0;  // grid points (grid runs)
    int64_t n_cells = 0;   // Widom cells
    double cost = 0.0;     // scheduling estimate per point
    double* host_out = nullptr;
    size_t tab_off = 0;
    std::vector<unsigned char> tab;
    // gradient runs: d LJ / d sigma_ts factors (Real) and mixing-rule Jacobian per (type, site)
    size_t gtab_off = 0, jac_off = 0;
    std::vector<unsigned char> gtab;
    std::vector<double4> jac;
    RunArgs ra{};
};

int widom_prepare(const pm_widom_run& in, PointRun& r) {
    pm_framework* fw = const_cast<pm_framework*>(in.framework);  // the image-list cache only
    const pm_forcefield* ff = in.forcefield;
    const pm_ewald* ew = in.ewald;
    const pm_probe* pr = in.probe;
    if (!fw || !ff || !pr || in.n_blocks < 1 || in.n_per_block < 1)
        return fail(PM_ERR_ARG, "pm_widom: null handle or empty block layout");
    if (!(in.temperature > 0.0) || !std::isfinite(in.temperature))
        return fail(PM_ERR_ARG, "pm_widom: temperature must be positive and finite");
    if (int rc = check_species(fw, ff)) return rc;
    for (int s = 0; s < pr->ns; ++s)
        if (pr->site[s].lj && pr->lj_species[s] >= ff->n)
            return fail(PM_ERR_ARG, "pm_widom: probe species outside the force field");
    const bool es = fw->charged && pr->charged;  // src/henry.jl:93
    if (es && !ew) return fail(PM_ERR_ARG, "pm_widom: charged framework and probe need an Ewald handle");
    const int ns = pr->ns;
    const int64_t chunks = (in.n_per_block + WG - 1) / WG;
    const int64_t n_wg = chunks * in.n_blocks;
    if (n_wg > (int64_t)INT32_MAX / 2 || (int64_t)in.n_blocks * in.n_per_block > ((int64_t)1 << 53))
        return fail(PM_ERR_ARG, "pm_widom: too many insertions for one run");
    r.fw = fw;
    r.ff = ff;
    r.ew = es ? ew : nullptr;
    r.pr = pr;
    r.es = es;
    r.ns = ns;
    r.n_wg = n_wg;
    RunArgs& a = r.ra;
    fill_run_pair(a, fw, ff, es ? ew : nullptr, es, ns, pr->site);
    a.seed = in.seed;
    a.n_per_block = in.n_per_block;
    a.rotate = pr->rotate;
    if (in.access) {
        a.acc = in.access->d;
        for (int d = 0; d < 3; ++d) a.acc_n[d] = in.access->n[d];
    }
    a.n_blocks = in.n_blocks;
    a.chunks_per_block = (int)chunks;
    a.temperature = in.temperature;
    for (int s = 0; s < ns; ++s) {
        a.off[s][0] = pr->site[s].ox;
        a.off[s][1] = pr->site[s].oy;
        a.off[s][2] = pr->site[s].oz;
        a.site_q[s] = pr->site[s].q;
    }
    const double pts = std::min((double)in.n_blocks * (double)in.n_per_block * ns, (double)MAX_POINT_SLOTS);
    choose_cells(fw, pts, a.ncell, &a.chunk_r, &a.zbins);
    r.n_cells = (int64_t)a.ncell[0] * a.ncell[1] * a.ncell[2];
    const double vol = std::fabs(fw->F[0] * fw->F[4] * fw->F[8]);
    r.cost = (double)fw->n_rep / vol * std::pow(a.r_cmax + a.chunk_r, 3);
    return PM_OK;
}

int grid_prepare(const pm_grid_run& in, PointRun& r) {
    pm_framework* fw = const_cast<pm_framework*>(in.framework);  // the image-list cache only
    const pm_forcefield* ff = in.forcefield;
    const pm_ewald* ew = in.ewald;
    const pm_probe* pr = in.probe;
    if (!fw || !pr || !in.out) return fail(PM_ERR_ARG, "pm_energy_grid: null handle or pointer");
    if (pr->ns != 1) return fail(PM_ERR_ARG, "pm_energy_grid: probe must have exactly one site");
    if (pr->site[0].ox != 0.0 || pr->site[0].oy != 0.0 || pr->site[0].oz != 0.0)
        return fail(PM_ERR_ARG, "pm_energy_grid: single-site probe must sit at its centre of mass");
    for (int i = 0; i < 3; ++i)
        if (in.n_pts[i] < 1 || in.n_pts[i] > 65536) return fail(PM_ERR_ARG, "pm_energy_grid: n_pts out of range");
    if (pr->site[0].lj) {
        if (!ff) return fail(PM_ERR_ARG, "pm_energy_grid: LJ probe needs a force field");
        if (int rc = check_species(fw, ff)) return rc;
        if (pr->lj_species[0] >= ff->n)
            return fail(PM_ERR_ARG, "pm_energy_grid: probe species outside the force field");
    }
    const bool es = fw->charged && pr->charged;
    if (es && !ew) return fail(PM_ERR_ARG, "pm_energy_grid: charged probe needs an Ewald handle");
    const int brick[3] = {8, 8, 4};
    int nbrick[3];
    for (int i = 0; i < 3; ++i) nbrick[i] = (in.n_pts[i] + brick[i] - 1) / brick[i];
    const long long n_bricks = (long long)nbrick[0] * nbrick[1] * nbrick[2];
    const long long n_pts = (long long)in.n_pts[0] * in.n_pts[1] * in.n_pts[2];
    if (n_bricks > INT32_MAX / 2 || n_pts > (1LL << 40))
        return fail(PM_ERR_ARG, "pm_energy_grid: grid too large");
    r.fw = fw;
    r.ff = ff;
    r.ew = es ? ew : nullptr;
    r.pr = pr;
    r.es = es;
    r.ns = 1;
    r.n_wg = n_bricks;
    r.n_points = n_pts;
    r.host_out = in.out;
    RunArgs& a = r.ra;
    fill_run_pair(a, fw, pr->site[0].lj ? ff : nullptr, es ? ew : nullptr, es, 1, pr->site);
    for (int i = 0; i < 3; ++i) {
        a.grid_n[i] = in.n_pts[i];
        a.brick[i] = brick[i];
        a.nbrick[i] = nbrick[i];
    }
    a.grid_q = pr->site[0].q;
    // Cartesian span of one brick and its half diagonal
    double ed[3][3];
    for (int d = 0; d < 3; ++d) {
        const double step =
            in.n_pts[d] > 1 ? (double)(brick[d] - 1) / (double)(in.n_pts[d] - 1) / fw->rep[d] : 0.0;
        for (int i = 0; i < 3; ++i) ed[d][i] = fw->F[3 * d + i] * step;
    }
    a.chunk_r = half_diagonal(ed[0], ed[1], ed[2]);
    const double vol = std::fabs(fw->F[0] * fw->F[4] * fw->F[8]);
    r.cost = (double)fw->n_rep / vol * std::pow(a.r_cmax + a.chunk_r, 3);
    return PM_OK;
}

template <typename Real, int NT>
void pair_launch_t(const GroupArgs& g, int n_wg, size_t shm) {
    hipLaunchKernelGGL((point_pair_kernel<Real, NT>), dim3(n_wg), dim3(WG), shm, g_ctx.stream, g);
}

// Compiled type counts of the gradient kernels: a gradient run of n_gt
// framework types runs the smallest one that holds it (its per-type FP64 sums
// take 6 NT registers per point).
int grad_nt(int n_gt) { return n_gt <= 2 ? 2 : n_gt <= 4 ? 4 : n_gt <= 6 ? 6 : GT; }

template <typename Real>
void pair_launch_grad(const GroupArgs& g, int n_wg, size_t shm, int nt) {
    switch (nt) {
        case 2: pair_launch_t<Real, 2>(g, n_wg, shm); break;
        case 4: pair_launch_t<Real, 4>(g, n_wg, shm); break;
        case 6: pair_launch_t<Real, 6>(g, n_wg, shm); break;
        default: pair_launch_t<Real, GT>(g, n_wg, shm); break;
    }
}

template <int NT>
void fourier_grad_launch_t(const GroupArgs& g, unsigned n_wg) {
    hipLaunchKernelGGL(point_fourier_grad_kernel<NT>, dim3(n_wg), dim3(FOURIER_WG), 0, g_ctx.stream, g);
}

template <int NT>
void combine_grad_launch_t(const GroupArgs& g, unsigned n_wg) {
    hipLaunchKernelGGL(widom_combine_grad_kernel<NT>, dim3(n_wg), dim3(WG), 0, g_ctx.stream, g);
}

// The gradient combine for the group's largest type count.
void combine_grad_launch(const GroupArgs& g, unsigned n_wg, int n_gt) {
    switch (n_gt) {
        case 1: combine_grad_launch_t<1>(g, n_wg); break;
        case 2: combine_grad_launch_t<2>(g, n_wg); break;
        case 3: combine_grad_launch_t<3>(g, n_wg); break;
        case 4: combine_grad_launch_t<4>(g, n_wg); break;
        case 5: combine_grad_launch_t<5>(g, n_wg); break;
        case 6: combine_grad_launch_t<6>(g, n_wg); break;
        case 7: combine_grad_launch_t<7>(g, n_wg); break;
        default: combine_grad_launch_t<GT>(g, n_wg); break;
    }
}

// The Fourier sums' work is proportional to the type count: one kernel per count.
void fourier_grad_launch(const GroupArgs& g, unsigned n_wg, int n_gt) {
    switch (n_gt) {
        case 1: fourier_grad_launch_t<1>(g, n_wg); break;
        case 2: fourier_grad_launch_t<2>(g, n_wg); break;
        case 3: fourier_grad_launch_t<3>(g, n_wg); break;
        case 4: fourier_grad_launch_t<4>(g, n_wg); break;
        case 5: fourier_grad_launch_t<5>(g, n_wg); break;
        case 6: fourier_grad_launch_t<6>(g, n_wg); break;
        case 7: fourier_grad_launch_t<7>(g, n_wg); break;
        default: fourier_grad_launch_t<GT>(g, n_wg); break;
    }
}

unsigned grid1d(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Device memory a call must use instead of the library's growable scratch
// (pm_widom_noalloc): the call then allocates nothing, and fails before any
// device work if its buffers do not fit or an image list is not built yet.
struct FixedScratch {
    void* base;
    size_t bytes;
    const char* caller;
};

// Runs every run of a call through the point engine.  Widom: out_sums gets
// n_blocks x (S, W_vdw, W_es) per run in input order; with grad (gradient
// runs, their gtab/jac filled) out_gsums gets n_blocks x 3 GT sums of
// w dE/d(eps_t, sigma_t, q_t).  Grids: each run's host_out gets its grid
// (i fastest).
int execute_points(std::vector<PointRun>& runs, int mode, int precision, double* out_sums, bool grad = false,
                   double* out_gsums = nullptr, const FixedScratch* fixed = nullptr) {
    const bool fp64 = precision == PM_FP64;
    const int n_runs = (int)runs.size();
    int n_gt = 0, nt_grad = 0;  // gradient runs: framework types, and the pair kernel's compiled count
    if (grad) {
        for (const auto& r : runs) n_gt = std::max(n_gt, r.ra.n_gt);
        nt_grad = grad_nt(n_gt);
    }
    // image lists and paths
    for (auto& r : runs) {
        RunArgs& a = r.ra;
        a.fast = 0;
        a.n_img = 0;
        a.img = nullptr;
        a.img_type = nullptr;
        a.tile_sph = nullptr;
        a.tile_rng = nullptr;
        a.n_tiles = 0;
        if (a.chunk_r <= g_ctx.chunk_r_max) {
            const pm_framework::Images* im = nullptr;
            int rc = PM_OK;
            if (fixed) {
                im = find_images(r.fw, a.r_cmax + 0.01);
                if (!im)
                    return fail(PM_ERR_STATE, std::string(fixed->caller) +
                                                  ": the framework's image list for this cutoff is not built yet "
                                                  "(one pm_widom call with the same framework and force field "
                                                  "builds it; building it allocates device memory)");
            } else {
                rc = get_images(r.fw, a.r_cmax + 0.01, &im);
            }
            if (rc == PM_OK) {
                a.fast = 1;
                a.img = im->d_img;
                a.img_type = im->d_type;
                a.n_img = (int)im->n;
                a.tile_sph = im->d_tile_sph;
                a.tile_rng = im->d_tile_rng;
                a.n_tiles = (int)im->n_tiles;
            } else if (rc != PM_ERR_ARG) {
                return rc;
            }
        }
    }
    // scheduling order: the most expensive points first (results do not depend on it)
    std::vector<int> order(n_runs);
    for (int i = 0; i < n_runs; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return runs[x].cost > runs[y].cost; });

    // segments and launch groups
    struct Group {
        size_t seg0 = 0;
        int n_seg = 0;
        long long n_wg = 0, n_slots = 0, n_cells = 0;
        long long n_gen = 0;  // Widom: generator workgroups (GEN_CHUNKS insertion workgroups each)
        bool es = false;
        int sb_shift = 0;    // Widom: super-bins of 2^sb_shift cells
        int zstride = 1;     // Widom: the sort's slabs per cell (the largest of the group's runs)
        long long n_sb = 0;
    };
    std::vector<SegArgs> segs;
    std::vector<Group> groups;
    Group cur;
    auto close_group = [&]() {
        if (cur.n_seg > 0) groups.push_back(cur);
        cur = Group();
        cur.seg0 = segs.size();
    };
    for (int oi : order) {
        PointRun& r = runs[oi];
        if (mode == MODE_WIDOM) {
            const long long per_wg = (long long)WG * r.ns;
            int64_t w = 0;
            while (w < r.n_wg) {
                const long long room = (MAX_POINT_SLOTS - cur.n_slots) / per_wg;
                if (room <= 0 || (cur.n_seg > 0 && cur.n_cells + r.n_cells > MAX_GROUP_CELLS)) {
                    close_group();
                    continue;
                }
                const int64_t take = std::min<int64_t>(room, r.n_wg - w);
                SegArgs s{};
                s.run = oi;
                s.w0 = (int)w;
                s.wg_base = (int)cur.n_wg;
                s.cell_base = (unsigned)cur.n_cells;
                s.gen_base = (int)cur.n_gen;
                s.slot_base = cur.n_slots;
                s.n_slots = take * per_wg;
                segs.push_back(s);
                cur.n_seg++;
                cur.n_wg += take;
                cur.n_gen += (take + GEN_CHUNKS - 1) / GEN_CHUNKS;
                cur.n_slots += take * per_wg;
                cur.n_cells += r.n_cells;
                cur.es = cur.es || r.es;
                w += take;
            }
        } else {
            const long long padded = (r.n_points + WG - 1) / WG * WG;
            if (cur.n_seg > 0 && (cur.n_slots + padded > MAX_POINT_SLOTS || cur.n_wg + r.n_wg > INT32_MAX / 2))
                close_group();
            SegArgs s{};
            s.run = oi;
            s.wg_base = (int)cur.n_wg;
            s.slot_base = cur.n_slots;
            s.n_slots = r.n_points;
            segs.push_back(s);
            cur.n_seg++;
            cur.n_wg += r.n_wg;
            cur.n_slots += padded;
            cur.es = cur.es || r.es;
        }
    }
    close_group();
    long long max_slots = 0, max_cells = 0, max_items = 0, max_sb = 0;
    for (Group& gr : groups) {
        max_slots = std::max(max_slots, gr.n_slots);
        max_cells = std::max(max_cells, gr.n_cells);
        max_items = std::max(max_items, gr.n_cells + (gr.n_slots + WG - 1) / WG + 1);
        if (mode != MODE_WIDOM) continue;
        // super-bins of about SB_POINTS points; the sort's (cell, slab) keys
        // of a super-bin (2^S cells x the group's largest slab count) within
        // SB_KEYS_MAX, and no segment's cells spanning more than SB_SPAN_MAX
        int zstride = 1;
        for (int si = 0; si < gr.n_seg; ++si) zstride = std::max(zstride, runs[segs[gr.seg0 + si].run].ra.zbins);
        int s_max = 0;
        while (((2LL << s_max) * zstride) <= SB_KEYS_MAX) ++s_max;
        const double per_cell = (double)gr.n_slots / (double)std::max(gr.n_cells, 1LL);
        int S = (int)std::lround(std::log2(std::max(1.0, SB_POINTS / std::max(per_cell, 1.0))));
        S = std::min(std::max(S, 0), s_max);
        for (int si = 0; si < gr.n_seg; ++si) {
            const SegArgs& sg = segs[gr.seg0 + si];
            const long long c0 = sg.cell_base, c1 = c0 + runs[sg.run].n_cells - 1;
            while (S < s_max && (c1 >> S) - (c0 >> S) + 1 > SB_SPAN_MAX) ++S;
            if ((c1 >> S) - (c0 >> S) + 1 > SB_SPAN_MAX) return fail(PM_ERR_ARG, "too many Widom cells in one run");
        }
        gr.zstride = zstride;
        gr.sb_shift = S;
        gr.n_sb = (gr.n_cells + (1LL << S) - 1) >> S;
        max_sb = std::max(max_sb, gr.n_sb);
    }

    // LJ tables (host), per-run output offsets
    PM_HIP(hipStreamSynchronize(g_ctx.stream));  // the staging buffer may still be read by a copy
    size_t tab_total = 0, lj_lds = 0;
    long long part_total = 0, blocks_total = 0, grid_total = 0;
    std::vector<long long> part_off(n_runs), grid_off(n_runs);
    for (int i = 0; i < n_runs; ++i) {
        PointRun& r = runs[i];
        if (fp64) build_ljtab<double>(r.fw, r.ff, r.ns, r.pr->site, r.pr->lj_species, r.tab);
        else build_ljtab<float>(r.fw, r.ff, r.ns, r.pr->site, r.pr->lj_species, r.tab);
        r.tab_off = tab_total;
        tab_total += align256(r.tab.size());
        lj_lds = std::max(lj_lds, r.tab.size() + (grad ? r.gtab.size() : 0));
        if (grad) {
            r.gtab_off = tab_total;
            tab_total += align256(r.gtab.size());
            r.jac_off = tab_total;
            tab_total += align256(r.jac.size() * sizeof(double4));
        }
        r.ra.pstride = grad ? GK : 3;
        if (mode == MODE_WIDOM) {
            part_off[i] = part_total;
            part_total += r.n_wg * r.ra.pstride;
            r.ra.block_base = (int)blocks_total;
            blocks_total += r.ra.n_blocks;
        } else {
            grid_off[i] = grid_total;
            grid_total += (r.n_points * (fp64 ? 8 : 4) + 255) / 256 * 256;
        }
    }
    if (blocks_total > INT32_MAX / 2) return fail(PM_ERR_ARG, "too many Widom blocks in one call");
    Arena ar;
    const size_t o_runs = ar.take((size_t)n_runs * sizeof(RunArgs));
    const size_t o_segs = ar.take(segs.size() * sizeof(SegArgs));
    const size_t o_tabs = ar.take(tab_total);
    const size_t stage_bytes = ar.used;
    const bool widom = mode == MODE_WIDOM;
    const size_t o_part = ar.take(widom ? (size_t)part_total * sizeof(double) : 0);
    // what the call downloads, in one piece: the candidate-pair counter, then
    // the block sums, then the block derivative sums
    const size_t sums_bytes = widom ? (size_t)blocks_total * 3 * sizeof(double) : 0;
    const size_t gsums_bytes = grad ? (size_t)blocks_total * 3 * GT * sizeof(double) : 0;
    const size_t down_bytes = sizeof(unsigned long long) + sums_bytes + gsums_bytes;
    const size_t o_down = ar.take(down_bytes);
    const size_t o_grid = ar.take(widom ? 0 : (size_t)grid_total);
    const size_t o_pos = ar.take(widom ? (size_t)max_slots * sizeof(double4) : 0);
    const size_t o_cell = ar.take(widom ? (size_t)max_slots * sizeof(uint2) : 0);
    const size_t o_order = ar.take(widom ? (size_t)max_slots * sizeof(unsigned) : 0);
    const size_t o_lj = ar.take((size_t)max_slots * sizeof(double));
    const size_t o_sr = ar.take((size_t)max_slots * sizeof(double));
    const size_t o_lr = ar.take((size_t)max_slots * sizeof(double));
    const size_t o_count = ar.take(widom ? (size_t)(max_cells + 1) * sizeof(unsigned) : 0);
    const size_t o_poff = ar.take(widom ? (size_t)(max_cells + 1) * sizeof(unsigned) : 0);
    const size_t o_sloc = ar.take(widom ? (size_t)(max_cells + 1) * sizeof(unsigned) : 0);
    const size_t o_sbc = ar.take(widom ? (size_t)(max_sb * SB_COPIES + 1) * sizeof(unsigned) : 0);
    const size_t o_sbo = ar.take(widom ? (size_t)(max_sb * SB_COPIES + 1) * sizeof(unsigned) : 0);
    const size_t o_sbs = ar.take(widom ? (size_t)(max_sb + 1) * sizeof(unsigned) : 0);
    const size_t o_sbso = ar.take(widom ? (size_t)(max_sb + 1) * sizeof(unsigned) : 0);
    const size_t o_stage = ar.take(widom ? (size_t)max_slots * sizeof(uint2) : 0);
    const size_t o_items = ar.take(widom ? (size_t)max_items * sizeof(uint4) : 0);
    const size_t o_nitems = ar.take(sizeof(unsigned) * 4);
    const size_t o_g = ar.take(grad ? (size_t)max_slots * GQ * GT * sizeof(double) : 0);
    // separable Fourier part of charged grids: phases, B and C of the largest run
    size_t gf_elems = 0;
    if (!widom)
        for (const PointRun& r : runs)
            if (r.es) {
                const RunArgs& a = r.ra;
                const size_t n3 = (size_t)a.grid_n[2], n2 = (size_t)a.grid_n[1];
                gf_elems = std::max(gf_elems, (size_t)a.grid_n[0] + n2 + n3 +
                                                  (size_t)(a.Ma + 1) * (2 * a.Mb + 1) * n3 + (size_t)(a.Ma + 1) * n2 * n3);
            }
    const size_t o_gf = ar.take(gf_elems * sizeof(double2));
    const size_t o_sums = o_down + sizeof(unsigned long long);
    const size_t o_gsums = o_sums + sums_bytes;
    unsigned char* base = nullptr;
    if (fixed) {
        if (ar.used > fixed->bytes)
            return fail(PM_ERR_ARG, std::string(fixed->caller) + ": the call needs " + std::to_string(ar.used) +
                                        " bytes of scratch, above the fixed maximum of " +
                                        std::to_string(fixed->bytes) + " (PM_NOALLOC_SCRATCH_BYTES)");
        base = static_cast<unsigned char*>(fixed->base);
    } else {
        if (int rc = ensure_scratch(ar.used)) return rc;
        base = static_cast<unsigned char*>(g_ctx.scratch);
    }

    // device pointers of each run, then one staged upload of runs, segments, tables
    for (int i = 0; i < n_runs; ++i) {
        RunArgs& a = runs[i].ra;
        a.ljtab = base + o_tabs + runs[i].tab_off;
        if (grad) {
            a.gtab = base + o_tabs + runs[i].gtab_off;
            a.jac = reinterpret_cast<const double4*>(base + o_tabs + runs[i].jac_off);
            a.gsums = reinterpret_cast<double*>(base + o_gsums) + (long long)a.block_base * 3 * GT;
        }
        if (widom) {
            a.partials = reinterpret_cast<double*>(base + o_part) + part_off[i];
        } else {
            a.grid_d = fp64 ? reinterpret_cast<double*>(base + o_grid + grid_off[i]) : nullptr;
            a.grid_f = fp64 ? nullptr : reinterpret_cast<float*>(base + o_grid + grid_off[i]);
        }
    }
    // pinned staging when it is there or may be allocated: upload at its
    // start, download after it (the upload is done before the kernels run)
    unsigned char* pin = nullptr;
    const size_t up_bytes = align256(stage_bytes);
    if (int rc = io_staging(up_bytes + down_bytes, !fixed, &pin)) return rc;
    std::vector<unsigned char>& host_stage = g_ctx.host_tab;
    if (pin == nullptr) host_stage.resize(stage_bytes);
    unsigned char* const stage = pin != nullptr ? pin : host_stage.data();
    unsigned char* const down = pin != nullptr ? pin + up_bytes : nullptr;
    std::memset(stage, 0, stage_bytes);
    for (int i = 0; i < n_runs; ++i) {
        std::memcpy(stage + o_runs + i * sizeof(RunArgs), &runs[i].ra, sizeof(RunArgs));
        if (!runs[i].tab.empty())
            std::memcpy(stage + o_tabs + runs[i].tab_off, runs[i].tab.data(), runs[i].tab.size());
        if (grad) {
            if (!runs[i].gtab.empty())
                std::memcpy(stage + o_tabs + runs[i].gtab_off, runs[i].gtab.data(), runs[i].gtab.size());
            if (!runs[i].jac.empty())
                std::memcpy(stage + o_tabs + runs[i].jac_off, runs[i].jac.data(), runs[i].jac.size() * sizeof(double4));
        }
    }
    if (!segs.empty()) std::memcpy(stage + o_segs, segs.data(), segs.size() * sizeof(SegArgs));
    PM_HIP(hipMemcpyAsync(base, stage, stage_bytes, hipMemcpyHostToDevice, g_ctx.stream));

    const RunArgs* d_runs = reinterpret_cast<const RunArgs*>(base + o_runs);
    const SegArgs* d_segs = reinterpret_cast<const SegArgs*>(base + o_segs);
    unsigned* d_count = reinterpret_cast<unsigned*>(base + o_count);
    unsigned* d_poff = reinterpret_cast<unsigned*>(base + o_poff);
    unsigned* d_sloc = reinterpret_cast<unsigned*>(base + o_sloc);
    unsigned* d_sbc = reinterpret_cast<unsigned*>(base + o_sbc);
    unsigned* d_sbo = reinterpret_cast<unsigned*>(base + o_sbo);
    unsigned* d_sbs = reinterpret_cast<unsigned*>(base + o_sbs);
    unsigned* d_sbso = reinterpret_cast<unsigned*>(base + o_sbso);
    uint2* d_stage = reinterpret_cast<uint2*>(base + o_stage);
    unsigned* d_nitems = reinterpret_cast<unsigned*>(base + o_nitems);
    unsigned long long* d_cand = reinterpret_cast<unsigned long long*>(base + o_down);
    double* d_sums = reinterpret_cast<double*>(base + o_sums);

    GroupArgs g{};
    g.runs = d_runs;
    g.mode = mode;
    g.pos = reinterpret_cast<double4*>(base + o_pos);
    g.cell = reinterpret_cast<uint2*>(base + o_cell);
    g.sb_count = d_sbc;
    g.point_off = d_poff;
    g.items = reinterpret_cast<const uint4*>(base + o_items);
    g.n_items = d_nitems;
    g.order = reinterpret_cast<const unsigned*>(base + o_order);
    g.out_lj = reinterpret_cast<double*>(base + o_lj);
    g.out_sr = reinterpret_cast<double*>(base + o_sr);
    g.out_lr = reinterpret_cast<double*>(base + o_lr);
    g.cand_pairs = d_cand;
    g.out_g = grad ? reinterpret_cast<double*>(base + o_g) : nullptr;
    g.g_stride = max_slots;
    const size_t shm = (fp64 ? pair_lds_fixed<double>() : pair_lds_fixed<float>()) + std::max<size_t>(lj_lds, 16);

    // the counter is cleared by a kernel (Widom: the first cell-count fill),
    // which a caller's graph capture records safely
    if (!widom)
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, g_ctx.stream, static_cast<unsigned*>(nullptr), 0LL,
                           0u, d_cand);
    PM_HIP(hipEventRecord(g_ctx.ev0, g_ctx.stream));
    float pair_ms = 0.f;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group& gr = groups[gi];
        g.segs = d_segs + gr.seg0;
        g.n_seg = gr.n_seg;
        long long pair_wg = gr.n_wg;
        if (widom) {
            const long long n_cells = gr.n_cells, n_sb = gr.n_sb;
            const int S = gr.sb_shift;
            g.sb_shift = S;
            const long long n_sbc = n_sb * SB_COPIES;
            hipLaunchKernelGGL(fill_u32_kernel, dim3(grid1d(n_sbc + 1, 256)), dim3(256), 0, g_ctx.stream, d_sbc,
                               n_sbc + 1, 0u, gi == 0 ? d_cand : nullptr);
            hipLaunchKernelGGL(widom_gen_kernel, dim3((unsigned)gr.n_gen), dim3(WG * GEN_CHUNKS), 0, g_ctx.stream, g);
            // offsets of the super-bins' counter copies; their total (the
            // group's points) also closes point_off
            hipLaunchKernelGGL(scan_excl_kernel, dim3(1), dim3(SCAN_WG), 0, g_ctx.stream, d_sbc, (int)n_sbc, d_sbo,
                               d_poff + n_cells);
            hipLaunchKernelGGL(stage_points_kernel, dim3(grid1d(gr.n_slots, 256)), dim3(256), 0, g_ctx.stream,
                               g.cell, gr.n_slots, d_sbo, S, d_stage);
            hipLaunchKernelGGL(cell_sort_kernel, dim3((unsigned)n_sb), dim3(SCAN_WG), 0, g_ctx.stream, d_stage, d_sbo,
                               S, gr.zstride, (int)n_cells, d_count, d_poff, d_sloc, d_sbs,
                               const_cast<unsigned*>(g.order));
            hipLaunchKernelGGL(scan_excl_kernel, dim3(1), dim3(SCAN_WG), 0, g_ctx.stream, d_sbs, (int)n_sb, d_sbso,
                               d_nitems);
            hipLaunchKernelGGL(slice_items_kernel, dim3(grid1d(n_cells, 256)), dim3(256), 0, g_ctx.stream, d_count,
                               d_sloc, d_sbso, S, g.segs, gr.n_seg, const_cast<uint4*>(g.items), (int)n_cells);
            pair_wg = n_cells + (gr.n_slots + WG - 1) / WG;  // >= the slices the scan counts
        }
        PM_HIP(hipEventRecord(g_ctx.evp0, g_ctx.stream));
        if (grad) {
            if (fp64) pair_launch_grad<double>(g, (int)pair_wg, shm, nt_grad);
            else pair_launch_grad<float>(g, (int)pair_wg, shm, nt_grad);
        } else {
            if (fp64) pair_launch_t<double, 0>(g, (int)pair_wg, shm);
            else pair_launch_t<float, 0>(g, (int)pair_wg, shm);
        }
        PM_HIP(hipEventRecord(g_ctx.evp1, g_ctx.stream));
        const long long slot_wg = (gr.n_slots + WG - 1) / WG;
        if (gr.es && widom) {
            hipLaunchKernelGGL(point_fourier_kernel, dim3((unsigned)slot_wg), dim3(FOURIER_WG), 0, g_ctx.stream, g);
        } else if (gr.es) {
            // grids: the separable form, run by run
            for (int si = 0; si < gr.n_seg; ++si) {
                const SegArgs& sg = segs[gr.seg0 + si];
                const PointRun& r = runs[sg.run];
                if (!r.es) continue;
                const RunArgs& a = r.ra;
                GridFourierArgs f{};
                f.A = a.A;
                f.Ma = a.Ma; f.Mb = a.Mb; f.Mc = a.Mc;
                for (int d = 0; d < 3; ++d) { f.n[d] = a.grid_n[d]; f.repd[d] = a.repd[d]; }
                f.slot_base = sg.slot_base;
                f.out_lr = g.out_lr;
                f.z = reinterpret_cast<double2*>(base + o_gf);
                f.B = f.z + (f.n[0] + f.n[1] + f.n[2]);
                f.C = f.B + (long long)(a.Ma + 1) * (2 * a.Mb + 1) * f.n[2];
                const long long n_z = (long long)f.n[0] + f.n[1] + f.n[2];
                const long long n_b = (long long)(a.Ma + 1) * (2 * a.Mb + 1) * f.n[2];
                const long long n_c = (long long)(a.Ma + 1) * f.n[1] * f.n[2];
                const long long n_p = (long long)f.n[0] * f.n[1] * f.n[2];
                hipLaunchKernelGGL(grid_phase_kernel, dim3(grid1d(n_z, 256)), dim3(256), 0, g_ctx.stream, f);
                hipLaunchKernelGGL(grid_fourier_c_kernel, dim3(grid1d(n_b, 256)), dim3(256), 0, g_ctx.stream, f);
                hipLaunchKernelGGL(grid_fourier_b_kernel, dim3(grid1d(n_c, 256)), dim3(256), 0, g_ctx.stream, f);
                hipLaunchKernelGGL(grid_fourier_a_kernel, dim3(grid1d(n_p, 256)), dim3(256), 0, g_ctx.stream, f);
            }
        }
        if (grad && gr.es) fourier_grad_launch(g, (unsigned)slot_wg, n_gt);
        if (widom && grad)
            combine_grad_launch(g, (unsigned)gr.n_wg, n_gt);
        else if (widom)
            hipLaunchKernelGGL(widom_combine_kernel, dim3((unsigned)gr.n_wg), dim3(WG), 0, g_ctx.stream, g);
        else hipLaunchKernelGGL(grid_combine_kernel, dim3((unsigned)slot_wg), dim3(WG), 0, g_ctx.stream, g);
        PM_HIP(hipGetLastError());
        if (gi + 1 < groups.size()) {
            // the next group reuses the events: read this group's pair-kernel time now
            PM_HIP(hipEventSynchronize(g_ctx.evp1));
            float ms = 0.f;
            PM_HIP(hipEventElapsedTime(&ms, g_ctx.evp0, g_ctx.evp1));
            pair_ms += ms;
        }
    }
    if (widom) {
        hipLaunchKernelGGL(block_sums_kernel, dim3((unsigned)blocks_total), dim3(WG), 0, g_ctx.stream, d_runs, n_runs,
                           d_sums);
        PM_HIP(hipGetLastError());
    }
    PM_HIP(hipEventRecord(g_ctx.ev1, g_ctx.stream));
    unsigned long long cand = 0;
    if (down != nullptr) {
        PM_HIP(hipMemcpyAsync(down, d_cand, down_bytes, hipMemcpyDeviceToHost, g_ctx.stream));
    } else {
        PM_HIP(hipMemcpyAsync(&cand, d_cand, sizeof(cand), hipMemcpyDeviceToHost, g_ctx.stream));
        if (widom)
            PM_HIP(hipMemcpyAsync(out_sums, d_sums, sums_bytes, hipMemcpyDeviceToHost, g_ctx.stream));
        if (grad)
            PM_HIP(hipMemcpyAsync(out_gsums, base + o_gsums, gsums_bytes, hipMemcpyDeviceToHost, g_ctx.stream));
    }
    std::vector<GridPiece> pieces;
    std::vector<std::vector<float>> tmp;
    double pairs = 0.0;
    if (widom) {
        for (const PointRun& r : runs) pairs += (double)r.ra.n_blocks * r.ra.n_per_block * r.ns * r.fw->n_rep;
    } else {
        // grids go to the callers' FP64 buffers in pieces through pinned
        // staging (FP32 device grids are widened on the host)
        const size_t es = fp64 ? sizeof(double) : sizeof(float);
        size_t total = 0;
        for (int i = 0; i < n_runs; ++i) {
            pairs += (double)runs[i].n_points * runs[i].fw->n_rep;
            total += (size_t)runs[i].n_points * es;
        }
        const bool staged = total <= MAX_PINNED;
        if (staged) {
            if (int rc = ensure_pinned(total)) return rc;
            const long long per = (long long)std::max(PIECE_BYTES, total / 256 + 1) / (long long)es;
            size_t pin_off = 0;
            for (int i = 0; i < n_runs; ++i) {
                const PointRun& r = runs[i];
                for (long long e0 = 0; e0 < r.n_points; e0 += per) {
                    const long long n = std::min(per, r.n_points - e0);
                    pieces.push_back(GridPiece{base + o_grid + grid_off[i] + (size_t)e0 * es,
                                               static_cast<unsigned char*>(g_ctx.pinned) + pin_off, r.host_out + e0,
                                               n});
                    pin_off += (size_t)n * es;
                }
            }
            if (int rc = download_grids(pieces, !fp64)) return rc;
        } else {
            // beyond the staging limit: straight to the callers' buffers (FP64
            // grids), or through a pageable buffer per grid (FP32 grids, widened
            // below)
            tmp.resize(fp64 ? 0 : n_runs);
            for (int i = 0; i < n_runs; ++i) {
                const PointRun& r = runs[i];
                if (fp64) {
                    PM_HIP(hipMemcpyAsync(r.host_out, base + o_grid + grid_off[i],
                                          (size_t)r.n_points * sizeof(double), hipMemcpyDeviceToHost, g_ctx.stream));
                } else {
                    tmp[i].resize((size_t)r.n_points);
                    PM_HIP(hipMemcpyAsync(tmp[i].data(), base + o_grid + grid_off[i],
                                          (size_t)r.n_points * sizeof(float), hipMemcpyDeviceToHost, g_ctx.stream));
                }
            }
        }
    }
    if (int rc = finish_timing(pairs)) return rc;
    if (down != nullptr) {
        std::memcpy(&cand, down, sizeof(cand));
        if (widom) std::memcpy(out_sums, down + sizeof(cand), sums_bytes);
        if (grad) std::memcpy(out_gsums, down + sizeof(cand) + sums_bytes, gsums_bytes);
    }
    for (size_t i = 0; i < tmp.size(); ++i)
        for (long long p = 0; p < runs[i].n_points; ++p) runs[i].host_out[p] = (double)tmp[i][(size_t)p];
    {
        float ms = 0.f;
        PM_HIP(hipEventElapsedTime(&ms, g_ctx.evp0, g_ctx.evp1));
        pair_ms += ms;
    }
    g_ctx.last_pair_ms = pair_ms;
    g_ctx.last_cand = (double)cand;
    return PM_OK;
}

// Per-type Fourier boxes of an Ewald handle (gradient runs), built once: the
// structure factor of each framework species with unit charges on every atom
// of that species, on the handle's kept k-vectors.
int ensure_type_boxes(pm_ewald* ew, const pm_framework* fw) {
    if (ew->d_box_t) return PM_OK;
    if ((int)fw->species_global.size() > GT)
        return fail(PM_ERR_ARG, "gradient runs support at most PM_MAX_GRAD_TYPES framework species");
    const size_t box_n = (size_t)(ew->Ma + 1) * (2 * ew->Mb + 1) * (2 * ew->Mc + 1);
    std::vector<double4> ua((size_t)std::max<int64_t>(fw->n_unit, 1));
    for (int64_t u = 0; u < fw->n_unit; ++u)
        ua[(size_t)u] = make_double4(fw->unit_xf[3 * u], fw->unit_xf[3 * u + 1], fw->unit_xf[3 * u + 2],
                                     (double)fw->unit_type[u]);
    double4* d_ua = nullptr;
    double2* d_bt = nullptr;
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    hipError_t e = hipMalloc(&d_bt, box_n * GT * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&d_ua, ua.size() * sizeof(double4));
    if (e == hipSuccess) e = hipMemsetAsync(d_bt, 0, box_n * GT * sizeof(double2), g_ctx.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_ua, ua.data(), ua.size() * sizeof(double4), hipMemcpyHostToDevice, g_ctx.stream);
    if (e == hipSuccess && ew->nk > 0 && fw->n_unit > 0) {
        hipLaunchKernelGGL(type_structure_factor_kernel, dim3((unsigned)ew->nk), dim3(64), 0, g_ctx.stream, d_ua,
                           (long long)fw->n_unit, ew->d_k, ew->Mb, ew->Mc, d_bt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(d_ua);
    if (e != hipSuccess) {
        (void)hipFree(d_bt);
        return fail(PM_ERR_DEVICE, std::string("per-type structure factors: ") + hipGetErrorString(e));
    }
    ew->d_box_t = d_bt;
    return PM_OK;
}

// Gradient tables of a prepared Widom run: per (compact type t, site s) the
// factor 4 eps_ts / sigma_ts of d LJ / d sigma_ts (in the pair precision) and
// the mixing-rule Jacobian from jac[(i * n + j) * 4 + k] (i framework
// species, j probe species, k: d eps_ij/d eps_i, d eps_ij/d sigma_i,
// d sigma_ij/d eps_i, d sigma_ij/d sigma_i).
void fill_gradient_tables(PointRun& r, const double* jac, bool fp64) {
    const pm_framework* fw = r.fw;
    const pm_forcefield* ff = r.ff;
    const int nt = (int)fw->species_global.size(), ns = r.ns, n = ff->n;
    r.jac.assign((size_t)std::max(nt, 1) * ns, make_double4(0.0, 0.0, 0.0, 0.0));
    std::vector<double> gc((size_t)std::max(nt, 1) * ns, 0.0);
    for (int t = 0; t < nt; ++t)
        for (int s = 0; s < ns; ++s) {
            if (!r.pr->site[s].lj) continue;
            const int i = fw->species_global[t], j = r.pr->lj_species[s];
            const size_t idx = (size_t)i * n + j;
            const double sg = std::sqrt(ff->sig2[idx]);
            gc[(size_t)t * ns + s] = sg > 0.0 ? 4.0 * ff->eps[idx] / sg : 0.0;
            const double* J = jac + 4 * idx;
            r.jac[(size_t)t * ns + s] = make_double4(J[0], J[1], J[2], J[3]);
        }
    if (fp64) {
        r.gtab.resize(gc.size() * sizeof(double));
        std::memcpy(r.gtab.data(), gc.data(), r.gtab.size());
    } else {
        std::vector<float> f(gc.begin(), gc.end());
        r.gtab.resize(f.size() * sizeof(float));
        std::memcpy(r.gtab.data(), f.data(), r.gtab.size());
    }
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
    PM_HIP(hipEventCreate(&g_ctx.evp0));
    PM_HIP(hipEventCreate(&g_ctx.evp1));
    int lds = 0;
    PM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    g_ctx.max_lds = (size_t)std::max(lds, 0);
    // this device's copy of the no-alloc scratch array (the library binds one
    // device per process, so one address)
    void* p = nullptr;
    PM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PM_ERR_DEVICE, "pm_init: no device address for the no-alloc scratch array");
    g_ctx.noalloc_scratch = p;
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
    widths(fw->F, fw->width);
    for (int i = 0; i < 3; ++i) fw->rep[i] = rep[i];
    fw->n_unit = n_atoms;
    fw->n_rep = n_atoms * nr;
    fw->unit_type.resize((size_t)n_atoms);
    if (n_atoms > 0) fw->unit_xf.assign(xf, xf + 3 * n_atoms);
    fw->unit_q.assign((size_t)n_atoms, 0.0);
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
        fw->unit_type[i] = found;
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(xf[3 * i + d])) {
                delete fw;
                return fail(PM_ERR_ARG, "pm_framework_create: non-finite coordinate");
            }
        if (charge) {
            if (!std::isfinite(charge[i])) {
                delete fw;
                return fail(PM_ERR_ARG, "pm_framework_create: non-finite charge");
            }
            fw->unit_q[i] = charge[i];
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
                    atoms[k] = make_double4((xf[3 * i] + 1.0 * ra) / rep[0],
                                            (xf[3 * i + 1] + 1.0 * rb) / rep[1],
                                            (xf[3 * i + 2] + 1.0 * rc) / rep[2], fw->unit_q[i]);
                    sp[k] = fw->unit_type[i];
                }
    for (int64_t i = 0; i < n_atoms; ++i) {
        double q = fw->unit_q[i];
        fw->net_charge += q;
        if (q != 0.0) uq.push_back(make_double4(xf[3 * i], xf[3 * i + 1], xf[3 * i + 2], q));
    }
    fw->n_unit_q = (int64_t)uq.size();
    fw->charged = fw->n_unit_q > 0;
    auto cleanup = [&]() {
        (void)hipFree(fw->d_atoms); (void)hipFree(fw->d_species); (void)hipFree(fw->d_unit_q);
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
    (void)hipFree(fw->d_atoms);
    (void)hipFree(fw->d_species);
    (void)hipFree(fw->d_unit_q);
    for (auto& im : fw->images) {
        (void)hipFree(im.d_img);
        (void)hipFree(im.d_type);
        (void)hipFree(im.d_tile_sph);
        (void)hipFree(im.d_tile_rng);
    }
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
    int Ma = 0, Mb = 0, Mc = 0;
    for (int64_t k = 0; k < n_kvecs; ++k) {
        int32_t na = kidx[3 * k], nb = kidx[3 * k + 1], nc = kidx[3 * k + 2];
        if (!std::isfinite(wt[k])) return fail(PM_ERR_ARG, "pm_ewald_create: non-finite weight");
        if (na < 0 || (na == 0 && nb < 0) || (na == 0 && nb == 0 && nc <= 0))
            return fail(PM_ERR_ARG, "pm_ewald_create: k-vectors must be the half-space set (ka >= 0, ...)");
        if (na % fw->rep[0] || nb % fw->rep[1] || nc % fw->rep[2]) continue;
        int ma = na / fw->rep[0], mb = nb / fw->rep[1], mc = nc / fw->rep[2];
        m.push_back(ma);
        m.push_back(mb);
        m.push_back(mc);
        w.push_back(wt[k]);
        Ma = std::max(Ma, ma);
        Mb = std::max(Mb, std::abs(mb));
        Mc = std::max(Mc, std::abs(mc));
    }
    if ((int64_t)(Ma + 1) * (2 * Mb + 1) * (2 * Mc + 1) > (1LL << 26))
        return fail(PM_ERR_ARG, "pm_ewald_create: k-vector box too large");
    auto* ew = new pm_ewald();
    ew->alpha = alpha;
    ew->sr_cutoff = sr_cutoff;
    ew->nk = (int64_t)w.size();
    ew->Ma = Ma; ew->Mb = Mb; ew->Mc = Mc;
    size_t box_n = (size_t)(Ma + 1) * (2 * Mb + 1) * (2 * Mc + 1);
    int* d_m = nullptr;
    double* d_w = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(ew->d_k); (void)hipFree(ew->d_box); (void)hipFree(ew->d_erfc); (void)hipFree(ew->d_rows);
        (void)hipFree(ew->d_erfc64);
        (void)hipFree(d_m); (void)hipFree(d_w);
        delete ew;
    };
    // Cubic Hermite segments of erfc(alpha r) on [0, sr_cutoff], h = sr_cutoff /
    // ERFC_SEG, in t = r/h - i: c0 + c1 t + c2 t^2 + c3 t^3 matches erfc and its
    // derivative at both ends (interpolation error <= h^4 max|f''''| / 384).
    std::vector<float4> etab(ERFC_SEG);
    {
        const double h = sr_cutoff / ERFC_SEG;
        const double dscale = -2.0 * alpha / std::sqrt(M_PI);
        for (int i = 0; i < ERFC_SEG; ++i) {
            const double r0 = i * h, r1 = (i + 1) * h;
            const double f0 = std::erfc(alpha * r0), f1 = std::erfc(alpha * r1);
            const double d0 = h * dscale * std::exp(-alpha * alpha * r0 * r0);
            const double d1 = h * dscale * std::exp(-alpha * alpha * r1 * r1);
            etab[i] = make_float4((float)f0, (float)d0, (float)(3.0 * (f1 - f0) - 2.0 * d0 - d1),
                                  (float)(2.0 * (f0 - f1) + d0 + d1));
        }
    }
    // FP64 path: degree-5 segments of f(r) = erfc(alpha r) on [0, sr_cutoff],
    // h = sr_cutoff / ERFC64_SEG, in t = r/h - i, from the Taylor expansion about
    // the segment's midpoint m (in u = t - 1/2: sum_k f^(k)(m) h^k u^k / k!),
    // re-expanded in powers of t.  f^(k)(r) = alpha^k erfc^(k)(alpha r) with
    // erfc^(k)(x) = -(2/sqrt(pi)) (-1)^(k-1) H_(k-1)(x) e^(-x^2) (Hermite
    // polynomials); truncation error <= (h/2)^6 max|f^(6)| / 720, below 1e-15
    // for alpha h <= 0.01 (alpha r_c is about 3 to 4 for Ewald precisions
    // 1e-6 to 1e-8).  Evaluated in long double, stored as (c0, c1), (c2, c3),
    // (c4, c5).
    std::vector<double2> etab64(3 * ERFC64_SEG);
    {
        const long double h = (long double)sr_cutoff / ERFC64_SEG, al = alpha;
        const long double two_rpi = 2.0L / std::sqrt((long double)M_PI);
        for (int i = 0; i < ERFC64_SEG; ++i) {
            const long double x = al * h * (i + 0.5L);
            long double H[5];  // H_n(x), n = 0..4: H_(n+1) = 2x H_n - 2n H_(n-1)
            H[0] = 1.0L;
            H[1] = 2.0L * x;
            for (int n = 1; n < 4; ++n) H[n + 1] = 2.0L * x * H[n] - 2.0L * n * H[n - 1];
            const long double g = std::exp(-x * x);
            long double d[6];  // f^(k)(m) h^k / k!
            d[0] = std::erfc(x);
            long double scale = 1.0L;  // (alpha h)^k / k!
            for (int k = 1; k <= 5; ++k) {
                scale *= al * h / k;
                d[k] = -two_rpi * ((k - 1) % 2 ? -1.0L : 1.0L) * H[k - 1] * g * scale;
            }
            long double c[6] = {0.0L, 0.0L, 0.0L, 0.0L, 0.0L, 0.0L};
            for (int k = 0; k <= 5; ++k) {
                long double binom = 1.0L, pw = 1.0L;  // C(k, j) (-1/2)^(k-j), from j = k down
                for (int j = k; j >= 0; --j) {
                    c[j] += d[k] * binom * pw;
                    binom = binom * j / (k - j + 1);
                    pw *= -0.5L;
                }
            }
            etab64[3 * i] = make_double2((double)c[0], (double)c[1]);
            etab64[3 * i + 1] = make_double2((double)c[2], (double)c[3]);
            etab64[3 * i + 2] = make_double2((double)c[4], (double)c[5]);
        }
    }
    // kept mc range of every (ma, mb) row of the box
    const size_t n_rows = (size_t)(Ma + 1) * (2 * Mb + 1);
    std::vector<int2> rows(n_rows, make_int2(1, 0));
    for (size_t k = 0; k < w.size(); ++k) {
        int2& r = rows[(size_t)m[3 * k] * (2 * Mb + 1) + (size_t)(m[3 * k + 1] + Mb)];
        const int mc = m[3 * k + 2];
        if (r.x > r.y) r = make_int2(mc, mc);
        else r = make_int2(std::min(r.x, mc), std::max(r.y, mc));
    }
    if (hipMalloc(&ew->d_rows, n_rows * sizeof(int2)) != hipSuccess ||
        hipMalloc(&ew->d_erfc, ERFC_SEG * sizeof(float4)) != hipSuccess ||
        hipMalloc(&ew->d_erfc64, 3 * ERFC64_SEG * sizeof(double2)) != hipSuccess ||
        hipMalloc(&ew->d_box, box_n * sizeof(double2)) != hipSuccess ||
        hipMalloc(&ew->d_k, std::max<size_t>(ew->nk, 1) * sizeof(KVec)) != hipSuccess ||
        hipMalloc(&d_m, std::max<size_t>(m.size(), 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_w, std::max<size_t>(w.size(), 1) * sizeof(double)) != hipSuccess) {
        cleanup();
        return fail(PM_ERR_OOM, "pm_ewald_create: device allocation failed");
    }
    hipError_t e = hipMemsetAsync(ew->d_box, 0, box_n * sizeof(double2), g_ctx.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ew->d_erfc, etab.data(), ERFC_SEG * sizeof(float4), hipMemcpyHostToDevice,
                           g_ctx.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ew->d_erfc64, etab64.data(), 3 * ERFC64_SEG * sizeof(double2), hipMemcpyHostToDevice,
                           g_ctx.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ew->d_rows, rows.data(), n_rows * sizeof(int2), hipMemcpyHostToDevice, g_ctx.stream);
    if (e == hipSuccess && ew->nk > 0) {
        e = hipMemcpyAsync(d_m, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, g_ctx.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, g_ctx.stream);
        double nr = (double)fw->rep[0] * fw->rep[1] * fw->rep[2];
        if (e == hipSuccess) {
            hipLaunchKernelGGL(structure_factor_kernel, dim3((unsigned)ew->nk), dim3(64), 0, g_ctx.stream,
                               fw->d_unit_q, (long long)fw->n_unit_q, d_m, d_w, nr, ew->d_k);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(fourier_box_kernel, dim3((unsigned)((ew->nk + 255) / 256)), dim3(256), 0,
                               g_ctx.stream, ew->d_k, (int)ew->nk, Mb, Mc, ew->d_box);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(d_m);
    (void)hipFree(d_w);
    d_m = nullptr;
    d_w = nullptr;
    if (e != hipSuccess) {
        std::string msg = std::string("pm_ewald_create: ") + hipGetErrorString(e);
        cleanup();
        return fail(PM_ERR_DEVICE, msg);
    }
    *out = ew;
    return PM_OK;
}

int pm_ewald_destroy(pm_ewald* ew) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ew) return PM_OK;
    (void)hipFree(ew->d_k);
    (void)hipFree(ew->d_box);
    (void)hipFree(ew->d_erfc);
    (void)hipFree(ew->d_erfc64);
    (void)hipFree(ew->d_box_t);
    (void)hipFree(ew->d_rows);
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
    *out = pr;
    return PM_OK;
}

int pm_probe_destroy(pm_probe* pr) {
    delete pr;
    return PM_OK;
}

int pm_access_create(const int32_t n_pts[3], const uint8_t* data, pm_access** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!n_pts || !data || !out) return fail(PM_ERR_ARG, "pm_access_create: null pointer");
    int64_t npts = 1;
    for (int d = 0; d < 3; ++d) {
        if (n_pts[d] < 1 || n_pts[d] > 65536) return fail(PM_ERR_ARG, "pm_access_create: n_pts out of [1, 65536]");
        npts *= n_pts[d];
    }
    if (npts > ((int64_t)1 << 31)) return fail(PM_ERR_ARG, "pm_access_create: more than 2^31 grid points");
    std::vector<unsigned char> h((size_t)npts);
    for (int64_t p = 0; p < npts; ++p) h[(size_t)p] = data[p] != 0 ? 1 : 0;
    auto* acc = new pm_access();
    for (int d = 0; d < 3; ++d) acc->n[d] = n_pts[d];
    if (hipMalloc(&acc->d, (size_t)npts) != hipSuccess) {
        delete acc;
        return fail(PM_ERR_OOM, "pm_access_create: device allocation failed");
    }
    if (hipMemcpy(acc->d, h.data(), (size_t)npts, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(acc->d);
        delete acc;
        return fail(PM_ERR_DEVICE, "pm_access_create: host-to-device copy failed");
    }
    *out = acc;
    return PM_OK;
}

int pm_access_destroy(pm_access* acc) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!acc) return PM_OK;
    (void)hipFree(acc->d);
    delete acc;
    return PM_OK;
}

int pm_block_pockets(const int32_t n_pts[3], const double* data, double energy_tol, int32_t block_pockets,
                     int64_t max_cycles, uint8_t* out_accessible, int32_t* out_segments, int64_t out_counts[3]) {
    if (!n_pts || !data || !out_accessible) return fail(PM_ERR_ARG, "pm_block_pockets: null pointer");
    int64_t npts = 1;
    for (int d = 0; d < 3; ++d) {
        if (n_pts[d] < 1 || n_pts[d] > 65536) return fail(PM_ERR_ARG, "pm_block_pockets: n_pts out of [1, 65536]");
        npts *= n_pts[d];
    }
    if (npts > ((int64_t)1 << 31)) return fail(PM_ERR_ARG, "pm_block_pockets: more than 2^31 grid points");
    if (block_pockets != 0 && block_pockets != 1) return fail(PM_ERR_ARG, "pm_block_pockets: block_pockets must be 0 or 1");
    if (max_cycles < 0) return fail(PM_ERR_ARG, "pm_block_pockets: negative max_cycles");
    const int64_t cap = max_cycles == 0 ? (int64_t)100000000 : max_cycles;
    try {
        std::string err = pmhost::block_pockets(n_pts, data, energy_tol, block_pockets == 1, cap, out_accessible,
                                                out_segments, out_counts);
        if (!err.empty()) return fail(PM_ERR_ARG, err);
    } catch (const std::bad_alloc&) {
        return fail(PM_ERR_OOM, "pm_block_pockets: host allocation failed");
    }
    return PM_OK;
}

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
    PoseArgs a{};
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
    PoseArgs a{};
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

int pm_pose_gradient(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, int32_t n_sites,
                     const int32_t* site_species, const double* site_charges, int64_t n_poses,
                     const double* site_xf, int32_t precision, double* out_e, double* out_grad) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !out_e || !out_grad || n_poses < 0 || n_sites < 1 || n_sites > PM_MAX_SITES || n_sites == 7 ||
        !site_species || !site_charges || (n_poses > 0 && !site_xf))
        return fail(PM_ERR_ARG, "pm_pose_gradient: bad handle, pointer or size (sites in 1..6 or 8)");
    bool any_lj = false, any_q = false;
    for (int s = 0; s < n_sites; ++s) {
        if (site_species[s] >= 0) any_lj = true;
        if (!std::isfinite(site_charges[s])) return fail(PM_ERR_ARG, "pm_pose_gradient: non-finite charge");
        if (site_charges[s] != 0.0) any_q = true;
    }
    if (any_lj) {
        if (!ff) return fail(PM_ERR_ARG, "pm_pose_gradient: LJ sites need a force field");
        if (int rc = check_species(fw, ff)) return rc;
        for (int s = 0; s < n_sites; ++s)
            if (site_species[s] >= ff->n)
                return fail(PM_ERR_ARG, "pm_pose_gradient: site species outside the force field");
    }
    const bool es = fw->charged && any_q;
    if (es && !ew) return fail(PM_ERR_ARG, "pm_pose_gradient: charged sites need an Ewald handle");
    if (n_poses == 0) return PM_OK;
    if (n_poses > (int64_t)INT32_MAX - WG) return fail(PM_ERR_ARG, "pm_pose_gradient: too many poses");
    PoseArgs a{};
    a.ns = n_sites;
    a.n_items = n_poses;
    std::vector<int32_t> sp(n_sites, 0);
    for (int s = 0; s < n_sites; ++s) {
        a.site[s].lj = site_species[s] >= 0 ? 1 : 0;
        a.site[s].q = site_charges[s];
        sp[s] = site_species[s] >= 0 ? site_species[s] : 0;
    }
    std::vector<double> out3;
    if (int rc = run_poses(fw, any_lj ? ff : nullptr, es ? ew : nullptr, a, sp.data(), precision, es, site_xf, out3,
                           out_grad))
        return rc;
    std::memcpy(out_e, out3.data(), out3.size() * sizeof(double));
    return PM_OK;
}

int pm_widom(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
             const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
             int64_t n_per_block, int32_t precision, double* out_block_sums,
             double* out_energies, const pm_access* access) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!out_block_sums) return fail(PM_ERR_ARG, "pm_widom: null output pointer");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    pm_widom_run in{};
    in.framework = fw;
    in.forcefield = ff;
    in.ewald = ew;
    in.probe = pr;
    in.temperature = temperature;
    in.seed = seed;
    in.n_per_block = n_per_block;
    in.n_blocks = n_blocks;
    in.access = access;
    std::vector<PointRun> runs(1);
    if (int rc = widom_prepare(in, runs[0])) return rc;
    double* d_energies = nullptr;
    const size_t e_bytes = (size_t)n_blocks * (size_t)n_per_block * 2 * sizeof(double);
    if (out_energies) {
        if (hipMalloc(&d_energies, e_bytes) != hipSuccess)
            return fail(PM_ERR_OOM, "pm_widom: energy buffer allocation failed");
        runs[0].ra.out_e = d_energies;
    }
    int rc = execute_points(runs, MODE_WIDOM, precision, out_block_sums);
    if (rc == PM_OK && out_energies) {
        hipError_t e = hipMemcpy(out_energies, d_energies, e_bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(PM_ERR_DEVICE, std::string("pm_widom: copy: ") + hipGetErrorString(e));
    }
    if (d_energies) {
        (void)hipStreamSynchronize(g_ctx.stream);
        (void)hipFree(d_energies);
    }
    return rc;
}

int pm_widom_noalloc(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                     const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
                     int64_t n_per_block, int32_t precision, double* out_block_sums,
                     const pm_access* access) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!out_block_sums) return fail(PM_ERR_ARG, "pm_widom_noalloc: null output pointer");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    if (!g_ctx.noalloc_scratch) return fail(PM_ERR_STATE, "pm_widom_noalloc: no scratch address for this device");
    pm_widom_run in{};
    in.framework = fw;
    in.forcefield = ff;
    in.ewald = ew;
    in.probe = pr;
    in.temperature = temperature;
    in.seed = seed;
    in.n_per_block = n_per_block;
    in.n_blocks = n_blocks;
    in.access = access;
    std::vector<PointRun> runs(1);
    if (int rc = widom_prepare(in, runs[0])) return rc;
    const FixedScratch fixed{g_ctx.noalloc_scratch, (size_t)PM_NOALLOC_SCRATCH_BYTES, "pm_widom_noalloc"};
    return execute_points(runs, MODE_WIDOM, precision, out_block_sums, false, nullptr, &fixed);
}

int pm_widom_gradient(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                      const pm_probe* pr, double temperature, uint64_t seed, int32_t n_blocks,
                      int64_t n_per_block, int32_t precision, const double* jac,
                      double* out_block_sums, double* out_block_grads, const pm_access* access) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!out_block_sums || !out_block_grads || !jac) return fail(PM_ERR_ARG, "pm_widom_gradient: null pointer");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    pm_widom_run in{};
    in.framework = fw;
    in.forcefield = ff;
    in.ewald = ew;
    in.probe = pr;
    in.temperature = temperature;
    in.seed = seed;
    in.n_per_block = n_per_block;
    in.n_blocks = n_blocks;
    in.access = access;
    std::vector<PointRun> runs(1);
    if (int rc = widom_prepare(in, runs[0])) return rc;
    PointRun& r = runs[0];
    const int nt = (int)r.fw->species_global.size();
    if (nt > GT) return fail(PM_ERR_ARG, "pm_widom_gradient: more than PM_MAX_GRAD_TYPES framework species");
    for (int64_t k = 0; k < (int64_t)ff->n * ff->n * 4; ++k)
        if (!std::isfinite(jac[k])) return fail(PM_ERR_ARG, "pm_widom_gradient: non-finite Jacobian entry");
    if (r.es) {
        if (int rc = ensure_type_boxes(const_cast<pm_ewald*>(r.ew), r.fw)) return rc;
        r.ra.A_t = r.ew->d_box_t;
    }
    r.ra.n_gt = nt;
    fill_gradient_tables(r, jac, precision == PM_FP64);
    std::vector<double> gs((size_t)n_blocks * 3 * GT, 0.0);
    if (int rc = execute_points(runs, MODE_WIDOM, precision, out_block_sums, true, gs.data())) return rc;
    for (int64_t k = 0; k < (int64_t)n_blocks * ff->n * 3; ++k) out_block_grads[k] = 0.0;
    for (int b = 0; b < n_blocks; ++b)
        for (int t = 0; t < nt; ++t)
            for (int q = 0; q < 3; ++q)
                out_block_grads[((size_t)b * ff->n + r.fw->species_global[t]) * 3 + q] =
                    gs[(size_t)b * 3 * GT + 3 * t + q];
    return PM_OK;
}

int pm_widom_batch(int32_t n_runs, const pm_widom_run* runs_in, int32_t precision, double* out_block_sums) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (n_runs < 1 || n_runs > PM_MAX_BATCH_RUNS || !runs_in || !out_block_sums)
        return fail(PM_ERR_ARG, "pm_widom_batch: null pointer or run count outside [1, PM_MAX_BATCH_RUNS]");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    std::vector<PointRun> runs((size_t)n_runs);
    for (int32_t i = 0; i < n_runs; ++i)
        if (int rc = widom_prepare(runs_in[i], runs[i]))
            return fail(rc, "pm_widom_batch: run " + std::to_string(i) + ": " + g_err);
    return execute_points(runs, MODE_WIDOM, precision, out_block_sums);
}

int pm_energy_grid(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew,
                   const pm_probe* pr, const int32_t n_pts[3], int32_t precision, double* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!n_pts) return fail(PM_ERR_ARG, "pm_energy_grid: null n_pts");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    pm_grid_run in{};
    in.framework = fw;
    in.forcefield = ff;
    in.ewald = ew;
    in.probe = pr;
    for (int i = 0; i < 3; ++i) in.n_pts[i] = n_pts[i];
    in.out = out;
    std::vector<PointRun> runs(1);
    if (int rc = grid_prepare(in, runs[0])) return rc;
    return execute_points(runs, MODE_GRID, precision, nullptr);
}

int pm_energy_grid_batch(int32_t n_runs, const pm_grid_run* runs_in, int32_t precision) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (n_runs < 1 || n_runs > PM_MAX_BATCH_RUNS || !runs_in)
        return fail(PM_ERR_ARG, "pm_energy_grid_batch: null pointer or run count outside [1, PM_MAX_BATCH_RUNS]");
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    std::vector<PointRun> runs((size_t)n_runs);
    for (int32_t i = 0; i < n_runs; ++i)
        if (int rc = grid_prepare(runs_in[i], runs[i]))
            return fail(rc, "pm_energy_grid_batch: grid " + std::to_string(i) + ": " + g_err);
    return execute_points(runs, MODE_GRID, precision, nullptr);
}

int pm_set_tuning(int32_t key, double value) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!std::isfinite(value)) return fail(PM_ERR_ARG, "pm_set_tuning: non-finite value");
    switch (key) {
        case PM_TUNE_CELL_POINTS:
            if (value < 16.0 || value > 65536.0) return fail(PM_ERR_ARG, "pm_set_tuning: cell points outside [16, 65536]");
            g_ctx.cell_points = value;
            return PM_OK;
        case PM_TUNE_CHUNK_R_MAX:
            if (value < 0.5 || value > 20.0) return fail(PM_ERR_ARG, "pm_set_tuning: chunk radius outside [0.5, 20] A");
            g_ctx.chunk_r_max = value;
            return PM_OK;
        default:
            return fail(PM_ERR_ARG, "pm_set_tuning: unknown key");
    }
}

int pm_get_tuning(int32_t key, double* value) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!value) return fail(PM_ERR_ARG, "pm_get_tuning: null pointer");
    switch (key) {
        case PM_TUNE_CELL_POINTS: *value = g_ctx.cell_points; return PM_OK;
        case PM_TUNE_CHUNK_R_MAX: *value = g_ctx.chunk_r_max; return PM_OK;
        default: return fail(PM_ERR_ARG, "pm_get_tuning: unknown key");
    }
}

int pm_autotune_widom(const pm_framework* fw, const pm_forcefield* ff, const pm_ewald* ew, const pm_probe* pr,
                      double temperature, uint64_t seed, int32_t n_blocks, int64_t n_per_block, int32_t precision,
                      double* out_cell_points, double* out_chunk_r_max, double* out_ms) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (precision != PM_FP64 && precision != PM_MIXED)
        return fail(PM_ERR_ARG, "precision must be PM_FP64 or PM_MIXED");
    pm_widom_run in{};
    in.framework = fw;
    in.forcefield = ff;
    in.ewald = ew;
    in.probe = pr;
    in.temperature = temperature;
    in.seed = seed;
    in.n_per_block = n_per_block;
    in.n_blocks = n_blocks;
    {
        std::vector<PointRun> probe_run(1);
        if (int rc = widom_prepare(in, probe_run[0])) return rc;  // validates the arguments
    }
    // the cube that sets a cell's cross-section (or, for sparse runs, the cell);
    // the timing is not monotonic in it, so the list is dense at 256-640
    const double cps[] = {64.0, 128.0, 192.0, 256.0, 320.0, 384.0, 448.0, 512.0, 576.0, 640.0, 768.0, 896.0, 1280.0};
    // the widest radius limit first: a narrower one must be more than 1% faster
    // to be kept, since a wider limit keeps sparser runs (other frameworks,
    // fewer insertions) on the image-list path
    const double crs[] = {6.0, 4.5, 3.0};
    const double keep_cp = g_ctx.cell_points, keep_cr = g_ctx.chunk_r_max;
    double best_ms = INFINITY, best_cp = keep_cp, best_cr = keep_cr;
    std::vector<double> sums((size_t)n_blocks * 3);
    // cell layouts already timed: (cells per axis, image-list path).  A
    // candidate giving the same layout runs the same launches, so it could only
    // win by timing noise and is not timed again (the wider radius limit, timed
    // first, stands).
    std::vector<std::array<int, 5>> seen;
    for (double cr : crs)
        for (double cp : cps) {
            g_ctx.cell_points = cp;
            g_ctx.chunk_r_max = cr;
            {
                std::vector<PointRun> runs(1);
                if (int rc = widom_prepare(in, runs[0])) {
                    g_ctx.cell_points = keep_cp;
                    g_ctx.chunk_r_max = keep_cr;
                    return rc;
                }
                const RunArgs& ra = runs[0].ra;
                const std::array<int, 5> layout{ra.ncell[0], ra.ncell[1], ra.ncell[2], ra.zbins,
                                                ra.chunk_r <= cr ? 1 : 0};
                if (std::find(seen.begin(), seen.end(), layout) != seen.end()) continue;
                seen.push_back(layout);
            }
            double ms = INFINITY;
            for (int rep = 0; rep < 2; ++rep) {
                std::vector<PointRun> runs(1);
                int rc = widom_prepare(in, runs[0]);
                if (rc == PM_OK) rc = execute_points(runs, MODE_WIDOM, precision, sums.data());
                if (rc != PM_OK) {
                    g_ctx.cell_points = keep_cp;
                    g_ctx.chunk_r_max = keep_cr;
                    return rc;
                }
                ms = std::min(ms, g_ctx.last_ms);
            }
            if (ms < (cr == best_cr ? 1.0 : 0.99) * best_ms) {
                best_ms = ms;
                best_cp = cp;
                best_cr = cr;
            }
        }
    g_ctx.cell_points = best_cp;
    g_ctx.chunk_r_max = best_cr;
    if (out_cell_points) *out_cell_points = best_cp;
    if (out_chunk_r_max) *out_chunk_r_max = best_cr;
    if (out_ms) *out_ms = best_ms;
    return PM_OK;
}

int pm_copy_probe(int64_t n_bytes, int32_t repeats, double* out_ms) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (n_bytes < 16 || n_bytes % 16 != 0 || n_bytes > (int64_t(1) << 34) || repeats < 1 || repeats > 100)
        return fail(PM_ERR_ARG, "pm_copy_probe: bytes must be a multiple of 16 in [16, 2^34], repeats in [1, 100]");
    double2* src = nullptr;
    double2* dst = nullptr;
    if (hipMalloc(&src, (size_t)n_bytes) != hipSuccess || hipMalloc(&dst, (size_t)n_bytes) != hipSuccess) {
        (void)hipFree(src);
        return fail(PM_ERR_OOM, "pm_copy_probe: allocation failed");
    }
    hipError_t e = hipMemsetAsync(src, 0, (size_t)n_bytes, g_ctx.stream);
    const long long n16 = n_bytes / 16;
    const unsigned grid = (unsigned)std::min<long long>((n16 + WG - 1) / WG, 65536);
    float best = INFINITY;
    for (int r = 0; r < repeats && e == hipSuccess; ++r) {
        e = hipEventRecord(g_ctx.ev0, g_ctx.stream);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(WG), 0, g_ctx.stream, src, dst, n16);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(g_ctx.ev1, g_ctx.stream);
        if (e == hipSuccess) e = hipEventSynchronize(g_ctx.ev1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
        best = std::min(best, ms);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return fail(PM_ERR_DEVICE, std::string("pm_copy_probe: ") + hipGetErrorString(e));
    if (out_ms) *out_ms = best;
    return PM_OK;
}

int pm_last_kernel_stats(double* kernel_ms, double* pair_evals) {
    if (kernel_ms) *kernel_ms = g_ctx.last_ms;
    if (pair_evals) *pair_evals = g_ctx.last_pairs;
    return PM_OK;
}

int pm_last_pair_kernel_stats(double* pair_kernel_ms, double* candidate_pairs) {
    if (pair_kernel_ms) *pair_kernel_ms = g_ctx.last_pair_ms;
    if (candidate_pairs) *candidate_pairs = g_ctx.last_cand;
    return PM_OK;
}


// ------------------------------------------------------------------ GCMC --
int pm_gcmc_destroy(pm_gcmc* gc) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc) return PM_OK;
    (void)hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(gc->d_params);
    (void)hipFree(gc->d_chains);
    (void)hipFree(gc->d_atoms);
    (void)hipFree(gc->d_fg);
    (void)hipFree(gc->d_gg);
    (void)hipFree(gc->d_sites);
    (void)hipFree(gc->d_com);
    (void)hipFree(gc->d_dens);
    (void)hipFree(gc->d_tot);
    (void)hipFree(gc->d_incomplete);
    (void)hipFree(gc->d_cell_start);
    (void)hipFree(gc->d_cell_atoms);
    delete gc;
    return PM_OK;
}

int pm_gcmc_create(const pm_framework* fw, const pm_forcefield* ff, int32_t n_species,
                   const pm_probe* const* templates, int32_t n_chains, const double* temperatures,
                   const double* fugacities, const pm_gcmc_settings* st, pm_gcmc** out) {
    std::unique_lock<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!fw || !ff || !templates || !temperatures || !fugacities || !st || !out)
        return fail(PM_ERR_ARG, "pm_gcmc_create: null handle or pointer");
    if (n_species < 1 || n_species > PM_MAX_GCMC_SPECIES)
        return fail(PM_ERR_ARG, "pm_gcmc_create: n_species outside [1, PM_MAX_GCMC_SPECIES]");
    if (n_chains < 1 || n_chains > 65536) return fail(PM_ERR_ARG, "pm_gcmc_create: n_chains outside [1, 65536]");
    if (st->n_burn_cycles < 0 || st->n_sample_cycles < PM_GCMC_BLOCKS)
        return fail(PM_ERR_ARG, "pm_gcmc_create: need n_burn_cycles >= 0 and n_sample_cycles >= 5");
    if (st->sample_frequency < 1 || st->snapshot_frequency < 1)
        return fail(PM_ERR_ARG, "pm_gcmc_create: sample and snapshot frequencies must be >= 1");
    if (st->capacity < 1 || st->capacity > (1 << 20))
        return fail(PM_ERR_ARG, "pm_gcmc_create: capacity outside [1, 2^20]");
    if (!(st->delta > 0.0) || !std::isfinite(st->delta))
        return fail(PM_ERR_ARG, "pm_gcmc_create: translation step must be positive");
    if (int rc = check_species(fw, ff)) return rc;
    for (int32_t c = 0; c < n_chains; ++c) {
        if (!(temperatures[c] > 0.0) || !std::isfinite(temperatures[c]))
            return fail(PM_ERR_ARG, "pm_gcmc_create: temperatures must be positive");
        for (int32_t s = 0; s < n_species; ++s) {
            const double f = fugacities[(size_t)c * n_species + s];
            if (!(f > 0.0) || !std::isfinite(f)) return fail(PM_ERR_ARG, "pm_gcmc_create: fugacities must be positive");
        }
    }
    int64_t n_dens = 0;
    if (st->density_species >= 0) {
        if (st->density_species >= n_species) return fail(PM_ERR_ARG, "pm_gcmc_create: density species out of range");
        n_dens = 1;
        for (int d = 0; d < 3; ++d) {
            if (st->density_n_pts[d] < 1 || st->density_n_pts[d] > 4096)
                return fail(PM_ERR_ARG, "pm_gcmc_create: density grid points outside [1, 4096]");
            n_dens *= st->density_n_pts[d];
        }
        if (n_dens * n_chains > ((int64_t)1 << 31))
            return fail(PM_ERR_ARG, "pm_gcmc_create: density grids too large");
    }
    GcmcParams hp{};
    std::vector<int32_t> ads_ff;  // adsorbate site types -> force-field species
    bool any_rot = false;
    for (int32_t s = 0; s < n_species; ++s) {
        const pm_probe* pr = templates[s];
        if (!pr) return fail(PM_ERR_ARG, "pm_gcmc_create: null template");
        int k = 0;
        for (int i = 0; i < pr->ns; ++i) {
            if (!pr->site[i].lj) continue;  // charges take no part in the reference's GCMC
            const int32_t sp = pr->lj_species[i];
            if (sp >= ff->n) return fail(PM_ERR_ARG, "pm_gcmc_create: template species outside the force field");
            int t = -1;
            for (size_t q = 0; q < ads_ff.size(); ++q)
                if (ads_ff[q] == sp) t = (int)q;
            if (t < 0) {
                t = (int)ads_ff.size();
                ads_ff.push_back(sp);
            }
            hp.site_type[s][k] = t;
            hp.off[s][k][0] = pr->site[i].ox;
            hp.off[s][k][1] = pr->site[i].oy;
            hp.off[s][k][2] = pr->site[i].oz;
            ++k;
        }
        if (k == 0) return fail(PM_ERR_ARG, "pm_gcmc_create: a template has no LJ sphere");
        hp.n_sites[s] = k;
        hp.rotatable[s] = pr->rotate;  // needs_rotations, src/mc_helpers.jl:8
        any_rot |= pr->rotate != 0;
    }
    if (st->density_species >= 0 && (st->density_site_mask >> hp.n_sites[st->density_species]) != 0u)
        return fail(PM_ERR_ARG, "pm_gcmc_create: density site mask names sites the template lacks");
    const int n_ads = (int)ads_ff.size();
    const int n_fw = (int)fw->species_global.size();
    std::vector<double2> fg((size_t)std::max(1, n_fw) * n_ads), gg((size_t)n_ads * n_ads);
    for (int i = 0; i < n_fw; ++i)
        for (int j = 0; j < n_ads; ++j) {
            const size_t q = (size_t)fw->species_global[i] * ff->n + ads_ff[j];
            fg[(size_t)i * n_ads + j] = make_double2(ff->sig2[q], ff->eps[q]);
        }
    for (int i = 0; i < n_ads; ++i)
        for (int j = 0; j < n_ads; ++j) {
            const size_t q = (size_t)ads_ff[i] * ff->n + ads_ff[j];
            gg[(size_t)i * n_ads + j] = make_double2(ff->sig2[q], ff->eps[q]);
        }
    // proposal probabilities (src/gcmc.jl:404-419), cumulative in move order
    double prob[PM_GCMC_MOVES] = {0.35, 0.35, 0.25, any_rot ? 0.25 : 0.0, 0.05, n_species > 1 ? 0.1 : 0.0};
    double tot = 0.0;
    for (double v : prob) tot += v;
    int last = 0;
    double cum = 0.0;
    for (int k = 0; k < PM_GCMC_MOVES; ++k) {
        prob[k] /= tot;
        cum += prob[k];
        hp.cum[k] = cum;
        if (prob[k] > 0.0) last = k;
    }
    for (int k = last; k < PM_GCMC_MOVES; ++k) hp.cum[k] = 2.0;
    for (int i = 0; i < 9; ++i) {
        hp.f[i] = fw->F[i];
        hp.c2f[i] = fw->c2f[i];
    }
    hp.r2_cut = ff->r2_cutoff;
    hp.omega = std::fabs(fw->F[0] * fw->F[4] * fw->F[8]);
    for (int d = 0; d < 3; ++d) hp.rep[d] = (double)fw->rep[d];
    hp.n_atoms = (int)fw->n_rep;
    hp.n_species = n_species;
    hp.n_ads_types = n_ads;
    hp.n_fw_types = std::max(1, n_fw);
    hp.total_sites = 0;
    for (int s = 0; s < PM_MAX_GCMC_SPECIES; ++s) {
        hp.site_prefix[s] = hp.total_sites;
        hp.total_sites += s < n_species ? hp.n_sites[s] : 0;
    }
    hp.capacity = st->capacity;
    hp.n_burn = st->n_burn_cycles;
    hp.n_sample = st->n_sample_cycles;
    hp.adjust_every = st->n_burn_cycles / 12;  // floor(n_burn / 12), src/gcmc.jl:635
    hp.cycles_per_block = st->n_sample_cycles / PM_GCMC_BLOCKS;
    hp.sample_frequency = st->sample_frequency;
    hp.snapshot_frequency = st->snapshot_frequency;
    hp.dens_species = st->density_species;
    hp.dens_sim_box = st->density_sim_box;
    hp.dens_mask = st->density_site_mask;
    for (int d = 0; d < 3; ++d) hp.dens_n[d] = st->density_species >= 0 ? st->density_n_pts[d] : 1;
    hp.seed = st->seed;
    // replicated framework atoms in Xtals order with their compact type
    std::vector<double4> atoms((size_t)std::max<int64_t>(1, fw->n_rep));
    {
        int64_t k = 0;
        for (int ra = 0; ra < fw->rep[0]; ++ra)
            for (int rb = 0; rb < fw->rep[1]; ++rb)
                for (int rc = 0; rc < fw->rep[2]; ++rc)
                    for (int64_t i = 0; i < fw->n_unit; ++i, ++k)
                        atoms[(size_t)k] = make_double4((fw->unit_xf[3 * i] + 1.0 * ra) / fw->rep[0],
                                                        (fw->unit_xf[3 * i + 1] + 1.0 * rb) / fw->rep[1],
                                                        (fw->unit_xf[3 * i + 2] + 1.0 * rc) / fw->rep[2],
                                                        (double)fw->unit_type[i]);
    }
    // cell lists for the guest-host sums: cells of about GC_CELL_EDGE over the
    // box, each listing the atoms whose nearest image lies within r_c + the
    // cell's half diagonal of its centre (so within r_c of any point of the
    // cell); kept when they are clearly shorter than the framework
    constexpr double GC_CELL_EDGE = 4.0;
    std::vector<int> cell_start, cell_atoms;
    int ncell[3] = {1, 1, 1};
    if (fw->n_rep > 0) {
        for (int d = 0; d < 3; ++d)
            ncell[d] = (int)std::max(1.0, std::min(64.0, std::floor(fw->width[d] / GC_CELL_EDGE)));
        const int nc = ncell[0] * ncell[1] * ncell[2];
        if ((double)nc * (double)fw->n_rep <= 4e8) {
            double ea[3], eb[3], ec[3];
            for (int i = 0; i < 3; ++i) {
                ea[i] = fw->F[i] / ncell[0];
                eb[i] = fw->F[3 + i] / ncell[1];
                ec[i] = fw->F[6 + i] / ncell[2];
            }
            const double R = std::sqrt(ff->r2_cutoff) + half_diagonal(ea, eb, ec) + 1e-6;
            const double R2 = R * R;
            const double* F = fw->F;
            cell_start.assign((size_t)nc + 1, 0);
            for (int c2 = 0; c2 < ncell[2]; ++c2)
                for (int c1 = 0; c1 < ncell[1]; ++c1)
                    for (int c0 = 0; c0 < ncell[0]; ++c0) {
                        const double cf[3] = {(c0 + 0.5) / ncell[0], (c1 + 0.5) / ncell[1], (c2 + 0.5) / ncell[2]};
                        for (size_t a = 0; a < atoms.size() && (int64_t)a < fw->n_rep; ++a) {
                            double d0 = atoms[a].x - cf[0], d1 = atoms[a].y - cf[1], d2 = atoms[a].z - cf[2];
                            d0 -= std::nearbyint(d0);
                            d1 -= std::nearbyint(d1);
                            d2 -= std::nearbyint(d2);
                            const double x = F[0] * d0 + F[3] * d1 + F[6] * d2;
                            const double y = F[4] * d1 + F[7] * d2;
                            const double z = F[8] * d2;
                            if (x * x + y * y + z * z <= R2) cell_atoms.push_back((int)a);
                        }
                        cell_start[(size_t)(c0 + ncell[0] * (c1 + ncell[1] * c2)) + 1] = (int)cell_atoms.size();
                    }
            if ((double)cell_atoms.size() > 0.6 * (double)nc * (double)fw->n_rep) {
                cell_start.clear();
                cell_atoms.clear();
            }
        }
    }
    for (int d = 0; d < 3; ++d) hp.n_cell[d] = ncell[d];
    auto* gc = new pm_gcmc();
    gc->n_chains = n_chains;
    gc->n_species = n_species;
    gc->capacity = st->capacity;
    gc->n_dens = n_dens;
    const size_t per_chain_sites = (size_t)n_species * st->capacity * PM_MAX_SITES;
    const size_t per_chain_com = (size_t)n_species * st->capacity;
    lk.unlock();  // pm_gcmc_destroy takes the lock on failure
    auto bail = [&](int code, const char* msg) {
        pm_gcmc_destroy(gc);
        return fail(code, msg);
    };
    lk.lock();
    if (hipMalloc(&gc->d_params, sizeof(GcmcParams)) != hipSuccess ||
        hipMalloc(&gc->d_chains, sizeof(GcmcChain) * n_chains) != hipSuccess ||
        hipMalloc(&gc->d_atoms, sizeof(double4) * atoms.size()) != hipSuccess ||
        hipMalloc(&gc->d_fg, sizeof(double2) * fg.size()) != hipSuccess ||
        hipMalloc(&gc->d_gg, sizeof(double2) * gg.size()) != hipSuccess ||
        hipMalloc(&gc->d_sites, sizeof(double4) * per_chain_sites * n_chains) != hipSuccess ||
        hipMalloc(&gc->d_com, sizeof(double4) * per_chain_com * n_chains) != hipSuccess ||
        hipMalloc(&gc->d_tot, sizeof(double) * 2 * n_chains) != hipSuccess ||
        hipMalloc(&gc->d_incomplete, sizeof(int)) != hipSuccess ||
        (!cell_start.empty() && hipMalloc(&gc->d_cell_start, sizeof(int) * cell_start.size()) != hipSuccess) ||
        (!cell_atoms.empty() && hipMalloc(&gc->d_cell_atoms, sizeof(int) * cell_atoms.size()) != hipSuccess) ||
        (n_dens > 0 && hipMalloc(&gc->d_dens, sizeof(unsigned) * (size_t)n_dens * n_chains) != hipSuccess)) {
        lk.unlock();
        return bail(PM_ERR_OOM, "pm_gcmc_create: device allocation failed");
    }
    hp.atoms = gc->d_atoms;
    hp.cell_start = cell_atoms.empty() ? nullptr : gc->d_cell_start;
    hp.cell_atoms = cell_atoms.empty() ? nullptr : gc->d_cell_atoms;
    hp.lj_fg = gc->d_fg;
    hp.lj_gg = gc->d_gg;
    gc->hp = hp;
    std::vector<GcmcChain> hc((size_t)n_chains);
    for (int32_t c = 0; c < n_chains; ++c) {
        GcmcChain& h = hc[(size_t)c];
        std::memset(&h, 0, sizeof(GcmcChain));
        h.temperature = temperatures[c];
        for (int32_t s = 0; s < n_species; ++s) h.fug[s] = fugacities[(size_t)c * n_species + s];
        h.delta = st->delta;
        h.sites = gc->d_sites + per_chain_sites * c;
        h.com = gc->d_com + per_chain_com * c;
        h.density = n_dens > 0 ? gc->d_dens + (size_t)n_dens * c : nullptr;
    }
    if (hipMemcpy(gc->d_params, &hp, sizeof(GcmcParams), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gc->d_chains, hc.data(), sizeof(GcmcChain) * n_chains, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gc->d_atoms, atoms.data(), sizeof(double4) * atoms.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gc->d_fg, fg.data(), sizeof(double2) * fg.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gc->d_gg, gg.data(), sizeof(double2) * gg.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (!cell_atoms.empty() &&
         (hipMemcpy(gc->d_cell_start, cell_start.data(), sizeof(int) * cell_start.size(), hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(gc->d_cell_atoms, cell_atoms.data(), sizeof(int) * cell_atoms.size(), hipMemcpyHostToDevice) !=
              hipSuccess)) ||
        (n_dens > 0 && hipMemset(gc->d_dens, 0, sizeof(unsigned) * (size_t)n_dens * n_chains) != hipSuccess)) {
        lk.unlock();
        return bail(PM_ERR_DEVICE, "pm_gcmc_create: host-to-device copy failed");
    }
    *out = gc;
    return PM_OK;
}

int pm_gcmc_set_molecules(pm_gcmc* gc, int32_t chain, int32_t species, int32_t n, const double* site_xf,
                          const double* com_xf) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc) return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: null handle");
    if (chain < 0 || chain >= gc->n_chains || species < 0 || species >= gc->n_species)
        return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: chain or species out of range");
    if (n < 0 || n > gc->capacity) return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: n outside [0, capacity]");
    if (n > 0 && (!site_xf || !com_xf)) return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: null coordinates");
    GcmcChain h;
    PM_HIP(hipMemcpy(&h, gc->d_chains + chain, sizeof(GcmcChain), hipMemcpyDeviceToHost));
    if (h.outer != 0 || h.t != 0) return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: the chain has already run");
    const int ns = gc->hp.n_sites[species];
    std::vector<double4> sites((size_t)std::max(1, n) * PM_MAX_SITES, make_double4(0.0, 0.0, 0.0, 0.0));
    std::vector<double4> com((size_t)std::max(1, n));
    for (int32_t m = 0; m < n; ++m) {
        for (int k = 0; k < ns; ++k) {
            const double* x = site_xf + ((size_t)m * ns + k) * 3;
            if (!std::isfinite(x[0]) || !std::isfinite(x[1]) || !std::isfinite(x[2]))
                return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: non-finite coordinate");
            sites[(size_t)m * PM_MAX_SITES + k] = make_double4(x[0], x[1], x[2], 0.0);
        }
        const double* c = com_xf + (size_t)m * 3;
        if (!(c[0] >= 0.0 && c[0] <= 1.0 && c[1] >= 0.0 && c[1] <= 1.0 && c[2] >= 0.0 && c[2] <= 1.0))
            return fail(PM_ERR_ARG, "pm_gcmc_set_molecules: a centre of mass lies outside the simulation box");
        com[(size_t)m] = make_double4(c[0], c[1], c[2], 0.0);
    }
    if (n > 0) {
        PM_HIP(hipMemcpy(h.sites + (size_t)species * gc->capacity * PM_MAX_SITES, sites.data(),
                         sizeof(double4) * (size_t)n * PM_MAX_SITES, hipMemcpyHostToDevice));
        PM_HIP(hipMemcpy(h.com + (size_t)species * gc->capacity, com.data(), sizeof(double4) * (size_t)n,
                         hipMemcpyHostToDevice));
    }
    PM_HIP(hipMemcpy((char*)(gc->d_chains + chain) + offsetof(GcmcChain, n) + sizeof(int) * species, &n, sizeof(int),
                     hipMemcpyHostToDevice));
    gc->energy_stale = true;
    return PM_OK;
}

namespace {

// System energies of every chain from scratch into out[2 c + {0, 1}].
int gcmc_total_energies(pm_gcmc* gc) {
    hipLaunchKernelGGL(gcmc_total_energy_kernel, dim3(gc->n_chains), dim3(GC_WG), 0, g_ctx.stream, gc->d_params,
                       gc->d_chains, gc->d_tot);
    PM_HIP(hipGetLastError());
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    return PM_OK;
}

}  // namespace

int pm_gcmc_run(pm_gcmc* gc, int64_t n_cycles) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!gc) return fail(PM_ERR_ARG, "pm_gcmc_run: null handle");
    if (n_cycles < 0) return fail(PM_ERR_ARG, "pm_gcmc_run: negative cycle count");
    if (gc->energy_stale) {
        // starting configurations: system energy of each chain (src/gcmc.jl:398-401)
        if (int rc = gcmc_total_energies(gc)) return rc;
        std::vector<double> tot((size_t)2 * gc->n_chains);
        PM_HIP(hipMemcpy(tot.data(), gc->d_tot, sizeof(double) * tot.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < gc->n_chains; ++c) {
            const double e[2] = {tot[2 * c], tot[2 * c + 1]};
            PM_HIP(hipMemcpy((char*)(gc->d_chains + c) + offsetof(GcmcChain, e_gh), e, sizeof(e),
                             hipMemcpyHostToDevice));
        }
        gc->energy_stale = false;
    }
    if (n_cycles == 0) return PM_OK;
    // dynamic LDS: the LJ tables (at most 48 KiB), then the framework when it
    // fits, then the chain's molecules when at least GC_MIN_LDS_MOLS per
    // species fit; static LDS and slack take the last 4 KiB
    constexpr int GC_MIN_LDS_MOLS = 16;
    const GcmcParams& hp = gc->hp;
    const size_t n_tab = (size_t)hp.n_fw_types * hp.n_ads_types + (size_t)hp.n_ads_types * hp.n_ads_types;
    const size_t lds_tables = sizeof(double4) * ((n_tab + 1) / 2);
    const size_t lds_atoms = sizeof(double4) * (size_t)hp.n_atoms;
    const size_t avail = g_ctx.max_lds > lds_tables + 4096 ? g_ctx.max_lds - lds_tables - 4096 : 0;
    const bool atoms_lds = hp.n_atoms > 0 && lds_atoms <= avail;
    size_t per_mol = 0;  // LDS bytes per unit of the molecule capacity
    for (int s = 0; s < hp.n_species; ++s) per_mol += sizeof(double4) * (size_t)(hp.n_sites[s] + 1);
    const size_t mol_room = avail - (atoms_lds ? lds_atoms : 0);
    int lds_cap = (int)std::min<size_t>((size_t)gc->capacity, mol_room / per_mol);
    if (gc->lds_mols_limit >= 0) lds_cap = std::min(lds_cap, gc->lds_mols_limit);
    // a smaller store only when the caller asked for one
    const int min_cap = gc->lds_mols_limit > 0 ? std::min(GC_MIN_LDS_MOLS, gc->lds_mols_limit) : GC_MIN_LDS_MOLS;
    const bool mols_lds = !gc->mols_in_memory && lds_cap >= min_cap;
    const size_t dyn = lds_tables + (atoms_lds ? lds_atoms : 0) + (mols_lds ? per_mol * (size_t)lds_cap : 0);
    auto launch = [&](bool a_lds, bool m_lds, int resume, size_t bytes) {
        const dim3 grid(gc->n_chains), block(GC_WG);
        const long long nc = (long long)n_cycles;
        if (a_lds && m_lds)
            hipLaunchKernelGGL((gcmc_kernel<true, true>), grid, block, bytes, g_ctx.stream, gc->d_params,
                               gc->d_chains, nc, resume, lds_cap, gc->d_incomplete);
        else if (a_lds)
            hipLaunchKernelGGL((gcmc_kernel<true, false>), grid, block, bytes, g_ctx.stream, gc->d_params,
                               gc->d_chains, nc, resume, 0, gc->d_incomplete);
        else if (m_lds)
            hipLaunchKernelGGL((gcmc_kernel<false, true>), grid, block, bytes, g_ctx.stream, gc->d_params,
                               gc->d_chains, nc, resume, lds_cap, gc->d_incomplete);
        else
            hipLaunchKernelGGL((gcmc_kernel<false, false>), grid, block, bytes, g_ctx.stream, gc->d_params,
                               gc->d_chains, nc, resume, 0, gc->d_incomplete);
    };
    (void)hipEventRecord(g_ctx.ev0, g_ctx.stream);
    if (mols_lds) {
        PM_HIP(hipMemsetAsync(gc->d_incomplete, 0, sizeof(int), g_ctx.stream));
        launch(atoms_lds, true, 0, dyn);
        PM_HIP(hipGetLastError());
        int incomplete = 0;
        PM_HIP(hipMemcpyAsync(&incomplete, gc->d_incomplete, sizeof(int), hipMemcpyDeviceToHost, g_ctx.stream));
        PM_HIP(hipStreamSynchronize(g_ctx.stream));
        if (incomplete > 0) {
            // the rest of the call, and later calls, with the molecules in device memory
            gc->mols_in_memory = true;
            launch(atoms_lds, false, 1, lds_tables + (atoms_lds ? lds_atoms : 0));
        }
    } else {
        launch(atoms_lds, false, 0, lds_tables + (atoms_lds ? lds_atoms : 0));
    }
    PM_HIP(hipGetLastError());
    (void)hipEventRecord(g_ctx.ev1, g_ctx.stream);
    PM_HIP(hipStreamSynchronize(g_ctx.stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
    g_ctx.last_ms = ms;
    return PM_OK;
}

int pm_gcmc_get_state(const pm_gcmc* gc, int32_t chain, pm_gcmc_state* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc || !out) return fail(PM_ERR_ARG, "pm_gcmc_get_state: null pointer");
    if (chain < 0 || chain >= gc->n_chains) return fail(PM_ERR_ARG, "pm_gcmc_get_state: chain out of range");
    GcmcChain h;
    PM_HIP(hipMemcpy(&h, gc->d_chains + chain, sizeof(GcmcChain), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(pm_gcmc_state));
    out->cycles_done = h.outer;
    out->markov_time = h.t;
    out->num_snapshots = h.num_snapshots;
    for (int s = 0; s < PM_MAX_GCMC_SPECIES; ++s) out->n[s] = h.n[s];
    out->block = h.block;
    out->error = h.error;
    out->delta = h.delta;
    out->energy_gh = h.e_gh;
    out->energy_gg = h.e_gg;
    for (int k = 0; k < PM_GCMC_MOVES; ++k) {
        out->n_proposed[k] = h.n_prop[k];
        out->n_accepted[k] = h.n_acc[k];
    }
    for (int b = 0; b < PM_GCMC_BLOCKS; ++b) {
        out->block_samples[b] = h.blk_samples[b];
        for (int s = 0; s < PM_MAX_GCMC_SPECIES; ++s) {
            out->block_n[b][s] = h.blk_n[b][s];
            out->block_n2[b][s] = h.blk_n2[b][s];
        }
        out->block_gh[b] = h.blk_gh[b];
        out->block_gg[b] = h.blk_gg[b];
        out->block_gh2[b] = h.blk_gh2[b];
        out->block_gg2[b] = h.blk_gg2[b];
        out->block_un[b] = h.blk_un[b];
    }
    return PM_OK;
}

int pm_gcmc_get_molecules(const pm_gcmc* gc, int32_t chain, int32_t species, int32_t max_n, int32_t* n,
                          double* site_xf, double* com_xf) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc || !n) return fail(PM_ERR_ARG, "pm_gcmc_get_molecules: null pointer");
    if (chain < 0 || chain >= gc->n_chains || species < 0 || species >= gc->n_species)
        return fail(PM_ERR_ARG, "pm_gcmc_get_molecules: chain or species out of range");
    GcmcChain h;
    PM_HIP(hipMemcpy(&h, gc->d_chains + chain, sizeof(GcmcChain), hipMemcpyDeviceToHost));
    *n = h.n[species];
    const int32_t m_copy = std::min(std::max(max_n, 0), h.n[species]);
    if (m_copy == 0) return PM_OK;
    if (!site_xf || !com_xf) return fail(PM_ERR_ARG, "pm_gcmc_get_molecules: null coordinates");
    const int ns = gc->hp.n_sites[species];
    std::vector<double4> sites((size_t)m_copy * PM_MAX_SITES), com((size_t)m_copy);
    PM_HIP(hipMemcpy(sites.data(), h.sites + (size_t)species * gc->capacity * PM_MAX_SITES,
                     sizeof(double4) * sites.size(), hipMemcpyDeviceToHost));
    PM_HIP(hipMemcpy(com.data(), h.com + (size_t)species * gc->capacity, sizeof(double4) * com.size(),
                     hipMemcpyDeviceToHost));
    for (int32_t m = 0; m < m_copy; ++m) {
        for (int k = 0; k < ns; ++k) {
            const double4 v = sites[(size_t)m * PM_MAX_SITES + k];
            double* x = site_xf + ((size_t)m * ns + k) * 3;
            x[0] = v.x;
            x[1] = v.y;
            x[2] = v.z;
        }
        com_xf[3 * m] = com[(size_t)m].x;
        com_xf[3 * m + 1] = com[(size_t)m].y;
        com_xf[3 * m + 2] = com[(size_t)m].z;
    }
    return PM_OK;
}

int pm_gcmc_total_energy(pm_gcmc* gc, int32_t chain, double* out_gh, double* out_gg) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    if (!gc || !out_gh || !out_gg) return fail(PM_ERR_ARG, "pm_gcmc_total_energy: null pointer");
    if (chain < 0 || chain >= gc->n_chains) return fail(PM_ERR_ARG, "pm_gcmc_total_energy: chain out of range");
    if (int rc = gcmc_total_energies(gc)) return rc;
    double e[2];
    PM_HIP(hipMemcpy(e, gc->d_tot + 2 * chain, sizeof(e), hipMemcpyDeviceToHost));
    *out_gh = e[0];
    *out_gg = e[1];
    return PM_OK;
}

int pm_gcmc_set_lds_molecules(pm_gcmc* gc, int32_t max_per_species) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc) return fail(PM_ERR_ARG, "pm_gcmc_set_lds_molecules: null handle");
    if (max_per_species < -1) return fail(PM_ERR_ARG, "pm_gcmc_set_lds_molecules: limit below -1");
    gc->lds_mols_limit = max_per_species;
    gc->mols_in_memory = false;
    return PM_OK;
}

int pm_gcmc_get_density(const pm_gcmc* gc, int32_t chain, double* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!gc || !out) return fail(PM_ERR_ARG, "pm_gcmc_get_density: null pointer");
    if (chain < 0 || chain >= gc->n_chains) return fail(PM_ERR_ARG, "pm_gcmc_get_density: chain out of range");
    if (gc->n_dens == 0) return fail(PM_ERR_ARG, "pm_gcmc_get_density: the handle tracks no density grid");
    std::vector<unsigned> h((size_t)gc->n_dens);
    PM_HIP(hipMemcpy(h.data(), gc->d_dens + (size_t)gc->n_dens * chain, sizeof(unsigned) * h.size(),
                     hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < gc->n_dens; ++i) out[i] = (double)h[(size_t)i];
    return PM_OK;
}

}  // extern "C"
