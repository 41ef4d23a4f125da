// phg.hip -- host side of the MI355X gas-phase pressure projection library:
// contexts, grids (device copies of the operator and multigrid hierarchy),
// batches of systems, the batched MG-PCG driver and the C ABI declared in
// include/photon_gas.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <limits>
#include <vector>

#include "../../include/photon_gas.h"
#include "device.hpp"
#include "hierarchy.hpp"
#include "kernels_pcg.hpp"
#include "kernels_step.hpp"
#include "kernels_transport.hpp"
#include "kernels_flow.hpp"
#include "kernels_leak.hpp"

// scalar transport: blocks aimed at per launch (tile columns x k-chunks):
// three resident blocks on each of the 256 CUs (its LDS allows three), and
// chunks down to one plane, since a block's planes are a serial chain
#ifndef PHG_TR_BLOCKS
#define PHG_TR_BLOCKS 768
#endif
#ifndef PHG_TR_KMIN
#define PHG_TR_KMIN 1
#endif

// Device workspace of phg_zone_mean_noalloc, the entry point that allocates
// nothing: the closed-zone partial sums (one per task) and then the means
// (one per closed zone) of the call in progress.  Each device holds its own
// copy until the process ends.  Host code never passes this name to a kernel
// (on the host it is only a stand-in); phg_create looks up the copy's device
// address with hipGetSymbolAddress and every use goes through that address.
constexpr int kNoAllocDoubles = 2048;   // 16 KiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace phg {

// ---------------------------------------------------------------------------
// profiling: HIP events around each launch, accumulated per kernel class
// ---------------------------------------------------------------------------
enum KClass {
    K_RESID_INIT, K_DIR_APPLY, K_PRE_RED, K_RB_BLACK, K_RESID0, K_POST_BLACK, K_POST_RED, K_FIN, K_VC_RESTRICT,
    K_VC_JAC, K_VC_RESID, K_VC_PROLONG, K_VC_COARSE, K_XFLUSH, K_ZONE, K_RHS, K_CORRECT, K_ENERGY, K_ADVECT,
    K_COUNT, K_PRE_SMOOTH, K_POST_SMOOTH, K_TURB, K_HEAT, K_MOMENTUM, K_COMBUST, K_GASMISC, K_LEAK, K_VC_DOWN,
    K_VC_UP, K_VC_COARSE_MFMA, K_NCLASS
};
static const char* kKName[K_NCLASS] = {"resid_init",  "dir_apply",  "pre_red",    "rb_black",  "resid0",
                                       "post_black",  "post_red",   "finalize",   "vc_restrict", "vc_jacobi",
                                       "vc_resid",    "vc_prolong", "vc_coarse",  "x_flush",   "zone_mean",
                                       "rhs",         "correct",    "energy",     "advect",    "count_active",
                                       "pre_smooth",  "post_smooth", "turbulence", "heat_diffusion",
                                       "momentum",    "combustion",  "gas_misc",   "leakage",   "vc_down",
                                       "vc_up",       "vc_coarse_mfma"};

struct GridDev {
    int nx = 0, ny = 0, nz = 0, n = 0;
    float dx = 0;
    int prec = PHG_PREC_FP32;
    HostHierarchy H;
    std::vector<int> zone;
    std::vector<uint8_t> zone_open;
    float *ax = nullptr, *ay = nullptr, *az = nullptr, *fx = nullptr, *fy = nullptr, *fz = nullptr;
    uint8_t* blocked = nullptr;
    void *axc = nullptr, *ayc = nullptr, *azc = nullptr;   // V-cycle coefficient copies
    struct Lev {
        int n = 0, ellk = 0;
        int* col = nullptr;
        void* val = nullptr;
        void* diag = nullptr;
        int* agg = nullptr;
        int* cptr = nullptr;
        int* cidx = nullptr;
        void* cv = nullptr;          // coarse levels under 65536 rows: ELL column and value packed
        uint32_t* cblk = nullptr;    // level 0: each level-1 aggregate as its 2x2x2 block
    };
    std::vector<Lev> lv;
    double* cinv = nullptr;       // the coarsest inverse whole (nc <= CP_MIN) ...
    double* cinvp = nullptr;      // ... or its upper triangle as 64x64 tiles
    int cnt = 0;                  // strips of cinvp (0: cinv)
    bool dense = false;           // the dense inverse exists (coarsest <= kDenseMax unknowns)
    int coarse = PHG_COARSE_EXACT;   // how batches created from now on solve the coarsest level
    int* gs_ord = nullptr;        // Gauss-Seidel level schedule of the coarsest level (HostHierarchy)
    int* gs_lptr = nullptr;
    float* invd = nullptr;        // 1/diag of level 0, 0 where diag <= 0
    uint8_t* mask = nullptr;      // level-0 face mask when every coefficient is 0 or 1
    uint32_t* aggm = nullptr;     // with it, each cell's level-1 aggregate | mask << 24 (one load for both)
    bool binary = false;
    int nclosed = 0, nztask = 0;
    int *zcells = nullptr, *zclosed = nullptr, *ztptr = nullptr;
    int4* ztask = nullptr;
    float* scratch_b = nullptr;   // RHS of phg_project
    float* vmax = nullptr;        // max reductions of the last call: maxVel, maxNu, maxDiv
    // the rest of the gas step
    uint8_t* lit_new = nullptr;   // combustion's new flame flags (N)
    double* cpart = nullptr;      // combustion's per-block heat release sums
    double* zsum = nullptr;       // closed zones: task sums, then zone sums
    float* zval = nullptr;        // closed zones: one value each (shift or pressure)
    float* wind = nullptr;        // [axis 0..1][side 0..1][k] boundary inflow (NaN: none)
    bool wind_set = false;
    float wind_key[4] = {0, 0, 0, 0};
    uint8_t *lx = nullptr, *ly = nullptr, *lz = nullptr;   // crack (leak) faces, or none
    // closed-zone leakage: zone k of the closed ones -> zone id, its volume
    int* czone = nullptr;
    float* zvol = nullptr;
    float* zp = nullptr;          // zone background pressures as passed in (every zone)
    float* zpn = nullptr;         // ... and their end-of-step values from the crack root search
    float* zdp = nullptr;         // closed zones' rise rates
    CrackDev crack;               // crack faces and the per-zone lists over them
    int crack_nb = 0;             // bisection entries
    float4* crack_pre = nullptr;  // per bisection entry: other side's pressure, densities, open area
    float* crack_q = nullptr;     // per crack: volume flow a -> b
    std::vector<void*> crack_allocs;
    // face updates: the set-up's generation (batches made before a rebuild
    // are stale), the wind-closed sides and the caller's coefficients there
    int gen = 0;
    bool has_wc = false;
    uint8_t wc[4] = {0, 0, 0, 0};
    std::vector<float> side_a[4];
    std::vector<void*> allocs;
    int single_batch = -1;
    DGrid dg() const {
        DGrid g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.n = n;
        g.ax = ax;
        g.ay = ay;
        g.az = az;
        g.fx = fx;
        g.fy = fy;
        g.fz = fz;
        g.blocked = blocked;
        return g;
    }
};

struct BatchDev {
    int nsys = 0, prec = 0, maxl = 0, maxnc = 0;
    bool binary = false;           // every grid has 0/1 coefficients: level-0 kernels read the face mask
    int maxnx = 0;                 // widest grid: sizes the LDS plane buffers of the level-0 kernels
    bool unfused = false;          // run the separate level-0 smoothing kernels (comparison runs)
    // Level-0 smoothing per system: the fused passes for 0/1 grids of
    // nx <= FUSE_MAX_NX and even nx * ny smoothed once, the separate half
    // sweeps for the others, each over its own tile tables
    int2* tabf = nullptr;          // fused systems: every column tile (TB-thread blocks)
    int tabf_len = 0;
    int2* tabwf = nullptr;         // fused systems: even column tiles (2 TB-thread blocks)
    int tabwf_len = 0;
    int2* tabqf = nullptr;         // fused systems on 4 TB-thread blocks: every fourth column tile
    int tabqf_len = 0;
    int2* tabu = nullptr;          // the other systems: every column tile
    int tabu_len = 0;
    int maxnx_f = 0;               // widest grid among the fused systems on TB or 2 TB threads
    int maxnu0_u = 0;              // most sweeps among the others
    size_t lds0() const { return 2 * (size_t)(TB + 2 * maxnx) * sizeof(float); }
    std::vector<int> grids;
    std::vector<int> gens;         // each grid's set-up generation when the batch was made
    std::vector<DSys> h_sys;
    DSys* d_sys = nullptr;
    std::vector<int2*> tab;        // level 0: (system, tile); tab[0] only
    int2* tabw = nullptr;          // level 0 entries of even column tiles (blocks of 2 TB columns)
    int tabw_len = 0;
    int fuse_nt = 0;               // threads per block of the fused passes (0: 2 TB up to nx = 128, 4 TB above)
    bool pair_dir = false;         // direction update on column pairs (k_dir_apply_pair)
    std::vector<int> tab_len;
    std::vector<int4*> tabg;       // levels >= 1: (first system, count, tile, levels | rows << 8) of same-grid groups
    std::vector<int> tabg_len;
    int2* tab_sys = nullptr;       // one entry per system
    int2* tab_z = nullptr;         // (system, zone task) for the closed-zone means
    int tab_z_len = 0;
    int4* tab_coarse = nullptr;    // (first system, count, row tile): coarsest-solve blocks
    int tab_coarse_len = 0;
    int4* tab_cm = nullptr;        // the same for the MFMA coarsest product (runs of >= kMfmaMinSys systems)
    int tab_cm_len = 0;
    int cg_sys = 1;                // systems per coarsest-solve block (64 KB of LDS)
    int4* tab_cp = nullptr;        // packed inverses: (first system, count, strip I, strip J >= I)
    int tab_cp_len = 0;
    int4* tab_cr = nullptr;        // packed inverses: (first system, count, strip)
    int tab_cr_len = 0;
    int cp_sys = 1;                // systems per packed-solve block
    int2* tab_gs = nullptr;        // systems whose coarsest level takes the Gauss-Seidel sweeps
    int tab_gs_len = 0;
    int maxnc_gs = 0;              // their largest coarsest level (LDS of k_vc_coarse_gs)
    // coarse levels in LDS: (system, first LDS level) of the systems whose levels
    // from that one down take k_vc_down / k_vc_up (one block each); tail_s per
    // system (nl: every coarse level as its own launches)
    int2* tab_tail = nullptr;
    int tab_tail_len = 0;
    size_t tail_lds = 0;
    std::vector<int> tail_s;
    double* scal = nullptr;
    int* state = nullptr;
    int* d_count = nullptr;
    int* h_count = nullptr;        // pinned
    std::vector<void*> allocs;
    hipGraph_t graph = nullptr;          // one CG iteration
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph_full = nullptr;     // set-up and graph_full_n iterations
    hipGraphExec_t graph_full_exec = nullptr;
    int graph_full_n = 0;
    int graph_full_hi = -1;              // the refine > 0 flag graph_full was captured with
};

}  // namespace phg

using namespace phg;

struct phg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int next_id = 1;
    std::map<int, GridDev*> grids;
    std::map<int, BatchDev*> batches;
    bool prof = false;
    unsigned long long prof_mask = ~0ull;   // kernel classes timed while profiling
    struct Pending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double prof_ms[K_NCLASS] = {};
    long long prof_n[K_NCLASS] = {};
    double* scratch_dev = nullptr;   // this device's copy of `scratch`
    size_t lds_block = 65536;        // LDS one block may take on this device
    int cus = 256;                   // its compute units
};

namespace {

int fail(phg_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

// Host set-up allocates with std::vector: an exception must not cross the C
// ABI, so the entry points that do host work run inside this guard
template <class F>
int guarded(phg_ctx* ctx, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(ctx, PHG_ENOMEM, "host memory exhausted");
    } catch (const std::exception& e) {
        return fail(ctx, PHG_EINVAL, "host exception: %s", e.what());
    }
}

#define HIP_OK(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail(ctx, PHG_EHIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                          \
    } while (0)

hipEvent_t take_event(phg_ctx* ctx) {
    if (!ctx->pool.empty()) {
        hipEvent_t e = ctx->pool.back();
        ctx->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void prof_drain(phg_ctx* ctx) {
    for (auto& p : ctx->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->prof_ms[p.cls] += ms;
            ctx->prof_n[p.cls] += 1;
        }
        ctx->pool.push_back(p.a);
        ctx->pool.push_back(p.b);
    }
    ctx->pending.clear();
}

template <typename... KArgs, typename... Args>
void launch_nt(phg_ctx* ctx, int cls, void (*kern)(KArgs...), int blocks, int threads, size_t smem, Args... args) {
    if (blocks <= 0) return;
    hipEvent_t a = nullptr, b = nullptr;
    const bool timed = ctx->prof && ((ctx->prof_mask >> cls) & 1ull);
    if (timed) {
        a = take_event(ctx);
        b = take_event(ctx);
        (void)hipEventRecord(a, ctx->stream);
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), smem, ctx->stream, args...);
    if (timed) {
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({cls, a, b});
    }
}
template <typename... KArgs, typename... Args>
void launch(phg_ctx* ctx, int cls, void (*kern)(KArgs...), int blocks, size_t smem, Args... args) {
    launch_nt(ctx, cls, kern, blocks, TB, smem, args...);
}

template <class T>
int dalloc(phg_ctx* ctx, std::vector<void*>& owner, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(ctx, PHG_ENOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    owner.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

template <class T>
int upload(phg_ctx* ctx, std::vector<void*>& owner, T** p, const T* src, size_t count) {
    int rc = dalloc(ctx, owner, p, count);
    if (rc) return rc;
    if (count) HIP_OK(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

uint16_t f32_to_f16_bits(float f) {
    _Float16 h = (_Float16)f;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}
uint16_t f32_to_bf16_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

// upload float data in the V-cycle coefficient type of the precision
int upload_ct(phg_ctx* ctx, std::vector<void*>& owner, void** p, const std::vector<float>& v, int prec) {
    if (prec == PHG_PREC_FP32) {
        float* q;
        int rc = upload(ctx, owner, &q, v.data(), v.size());
        *p = q;
        return rc;
    }
    std::vector<uint16_t> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = f32_to_f16_bits(v[i]);
    uint16_t* q;
    int rc = upload(ctx, owner, &q, h.data(), h.size());
    *p = q;
    return rc;
}

size_t vt_size(int prec) { return prec == PHG_PREC_FP32 ? 4 : 2; }

void free_all(std::vector<void*>& v) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

// tile table of one level, XCD-ordered: block b runs on XCD b % 8, and gets
// the (b / 8)-th entry of that XCD's contiguous share of the logical list
template <class E>
std::vector<E> xcd_order(const std::vector<E>& logical) {
    const int M = (int)logical.size();
    std::vector<E> phys(M);
    const int q = M / 8, r = M % 8;
    for (int b = 0; b < M; ++b) {
        const int x = b % 8, idx = b / 8;
        const int start = x * q + std::min(x, r);
        phys[b] = logical[start + idx];
    }
    return phys;
}

int tiles(int n) { return (n + TB - 1) / TB; }

// largest coarsest level solved densely: above the reference's 2000-unknown
// stopping size, so the hierarchy keeps the reference's levels whenever it
// coarsens at all; the LDS of a coarsest-solve block holds 8 such vectors
constexpr int kDenseMax = 2048;
// cells per block of the closed-zone sums
constexpr int kZoneChunk = 8192;
// level-0 blocks a launch should offer the chip (256 CUs)
constexpr int kLevel0Blocks = 4096;
// batches whose 256-column tiles number at least this many (two per CU) split
// their columns into chunks of at least kMinChunk planes only
constexpr int kColTilesFull = 512;
constexpr int kMinChunk = 32;
// whole-inverse coarsest-solve blocks below which its groups are made smaller
constexpr int kCoarseBlocks = 256;
// systems on one grid from which the coarsest product runs on the matrix cores,
// and the 16-system tiles of an MFMA block
constexpr int kMfmaMinSys = 32;
constexpr int kMfmaTiles = 4;
// most level-0 red-black sweeps per side of the V-cycle
constexpr int kMaxSweeps0 = 4;

// ---------------------------------------------------------------------------
// the V-cycle and one CG iteration as launch sequences
// ---------------------------------------------------------------------------
// C0: level-0 coefficient type of the V-cycle (FP32/FP16 faces or the 0/1
// mask), CC: coarse operator storage, VT: V-cycle vector storage
template <class C0, class CC, class VT, class ZT>
void vcycle_t(phg_ctx* ctx, BatchDev* B, int first) {
    const DSys* S = B->d_sys;
    // 0/1 grids of moderate width and even plane size smoothed once take the
    // two fused passes; the other systems the separate half sweeps
    const bool anyf = std::is_same<C0, Mask01>::value && B->tabf_len > 0;
    const int2* TU = B->tabu;
    const int nu = B->tabu_len;
    if (anyf) {
        // pre-smoothing and residual in one pass.  Column pairs per thread
        // for the widest grid: staging NT/2 + 2 nx, smoothing up to
        // NT/2 + nx + 1; 2 TB-thread blocks take the even column tiles.
        static_assert(TB == 256, "pair counts below assume 256 threads");
        // the LDS rings are sized for each configuration's widest grid
        auto go = [&](auto k1, auto k0, int nt, int p2, int p1) {
            const size_t lds = fuse_lds(nt, p2, p1, false);
            if (nt == 2 * TB)
                launch_nt(ctx, K_PRE_SMOOTH, first ? k1 : k0, B->tabwf_len, nt, lds, S, (const int2*)B->tabwf);
            else
                launch(ctx, K_PRE_SMOOTH, first ? k1 : k0, B->tabf_len, lds, S, (const int2*)B->tabf);
        };
        const int mx = B->maxnx_f;
        if (B->fuse_nt != TB) {
            if (B->tabwf_len > 0) {
                if (mx <= 128)
                    go(k_pre_fused<VT, ZT, true, 2 * TB, 1, 1>, k_pre_fused<VT, ZT, false, 2 * TB, 1, 1>, 2 * TB, 1, 1);
                else
                    go(k_pre_fused<VT, ZT, true, 2 * TB, 2, 1>, k_pre_fused<VT, ZT, false, 2 * TB, 2, 1>, 2 * TB, 2, 1);
            }
            if (B->tabqf_len)
                launch_nt(ctx, K_PRE_SMOOTH,
                          first ? k_pre_fused<VT, ZT, true, 4 * TB, 1, 1> : k_pre_fused<VT, ZT, false, 4 * TB, 1, 1>,
                          B->tabqf_len, 4 * TB, fuse_lds(4 * TB, 1, 1, false), S, (const int2*)B->tabqf);
        } else {
            if (mx <= 64)
                go(k_pre_fused<VT, ZT, true, TB, 1, 1>, k_pre_fused<VT, ZT, false, TB, 1, 1>, TB, 1, 1);
            else if (mx <= 127)
                go(k_pre_fused<VT, ZT, true, TB, 2, 1>, k_pre_fused<VT, ZT, false, TB, 2, 1>, TB, 2, 1);
            else
                go(k_pre_fused<VT, ZT, true, TB, 2, 2>, k_pre_fused<VT, ZT, false, TB, 2, 2>, TB, 2, 2);
        }
    }
    if (nu > 0) {
        // red from zero, black, then (red, black) for each further sweep of
        // the systems that make it, then the residual
        launch(ctx, K_PRE_RED, k_pre_red<C0, ZT>, nu, 0, S, TU, first);
        launch(ctx, K_RB_BLACK, k_rb_black<C0, ZT>, nu, B->lds0(), S, TU, 0);
        for (int v = 1; v < B->maxnu0_u; ++v) {
            launch(ctx, K_POST_RED, k_post_red<C0, ZT>, nu, B->lds0(), S, TU, v);
            launch(ctx, K_RB_BLACK, k_rb_black<C0, ZT>, nu, B->lds0(), S, TU, v);
        }
        launch(ctx, K_RESID0, k_resid0<C0, ZT, VT>, nu, B->lds0(), S, TU);
    }
    const std::vector<int4*>& TG = B->tabg;
    const std::vector<int>& NG = B->tabg_len;
    launch(ctx, K_VC_RESTRICT, k_vc_restrict<CC, VT>, NG[1], 0, S, (const int*)B->state, TG[1], 0);
    for (int l = 1; l + 1 < B->maxl; ++l) {
        launch(ctx, K_VC_JAC, k_vc_jac<CC, VT>, NG[l], 0, S, (const int*)B->state, TG[l], l, 1);
        launch(ctx, K_VC_RESID, k_vc_resid<CC, VT>, NG[l], 0, S, (const int*)B->state, TG[l], l);
        launch(ctx, K_VC_RESTRICT, k_vc_restrict<CC, VT>, NG[l + 1], 0, S, (const int*)B->state, TG[l + 1], l);
    }
    if (B->tab_tail_len > 0)
        launch_nt(ctx, K_VC_DOWN, k_vc_down<CC, VT>, B->tab_tail_len, CT_NT, B->tail_lds, S, (const int*)B->state,
                  (const int2*)B->tab_tail);
    auto coarse = [&](auto kern, int cgs) {
        launch(ctx, K_VC_COARSE, kern, B->tab_coarse_len, coarse_lds(cgs, B->maxnc), S, (const int*)B->state,
               (const int4*)B->tab_coarse);
    };
    if (B->tab_cm_len > 0)
        launch(ctx, K_VC_COARSE_MFMA, k_vc_coarse_mfma<VT, kMfmaTiles>, B->tab_cm_len, 0, S, (const int*)B->state,
               (const int4*)B->tab_cm);
    switch (B->cg_sys) {
        case 1: coarse(k_vc_coarse<VT, 1>, 1); break;
        case 4: coarse(k_vc_coarse<VT, 4>, 4); break;
        case CG_SYS_MAX / 2: coarse(k_vc_coarse<VT, CG_SYS_MAX / 2>, CG_SYS_MAX / 2); break;
        default: coarse(k_vc_coarse<VT, CG_SYS_MAX>, CG_SYS_MAX); break;
    }
    if (B->tab_gs_len > 0)
        launch(ctx, K_VC_COARSE, k_vc_coarse_gs<CC, VT>, B->tab_gs_len, 2 * (size_t)B->maxnc_gs * sizeof(float), S,
               (const int*)B->state, (const int2*)B->tab_gs);
    if (B->tab_cp_len > 0) {
        auto packed = [&](auto kern, int cgs) {
            launch(ctx, K_VC_COARSE, kern, B->tab_cp_len, coarse_p_lds(cgs), S, (const int*)B->state,
                   (const int4*)B->tab_cp);
        };
        switch (B->cp_sys) {
            case 1: packed(k_vc_coarse_p<VT, 1>, 1); break;
            case 4: packed(k_vc_coarse_p<VT, 4>, 4); break;
            default: packed(k_vc_coarse_p<VT, 8>, 8); break;
        }
        launch(ctx, K_VC_COARSE, k_vc_coarse_r<VT>, B->tab_cr_len, 0, S, (const int*)B->state, (const int4*)B->tab_cr);
    }
    if (B->tab_tail_len > 0)
        launch_nt(ctx, K_VC_UP, k_vc_up<CC, VT>, B->tab_tail_len, CT_NT, B->tail_lds, S, (const int*)B->state,
                  (const int2*)B->tab_tail);
    for (int l = B->maxl - 2; l >= 1; --l) {
        launch(ctx, K_VC_PROLONG, k_vc_prolong<VT>, NG[l], 0, S, (const int*)B->state, TG[l], l);
        launch(ctx, K_VC_JAC, k_vc_jac<CC, VT>, NG[l], 0, S, (const int*)B->state, TG[l], l, 0);
        launch(ctx, K_VC_JAC, k_vc_jac<CC, VT>, NG[l], 0, S, (const int*)B->state, TG[l], l, 1);
    }
    // post-smoothing mirrors the pre-smoothing (black, red per sweep), so M
    // stays symmetric; the last red sweep of each system leaves its r.z partials
    if (anyf) {
        auto go = [&](auto kern, int nt, int p2, int p1) {
            const size_t lds = fuse_lds(nt, p2, p1, true);
            if (nt == 2 * TB)
                launch_nt(ctx, K_POST_SMOOTH, kern, B->tabwf_len, nt, lds, S, (const int2*)B->tabwf);
            else
                launch(ctx, K_POST_SMOOTH, kern, B->tabf_len, lds, S, (const int2*)B->tabf);
        };
        const int mx = B->maxnx_f;
        if (B->fuse_nt != TB) {
            if (B->tabwf_len > 0) {
                if (mx <= 128)
                    go(k_post_fused<VT, ZT, 2 * TB, 1, 1>, 2 * TB, 1, 1);
                else
                    go(k_post_fused<VT, ZT, 2 * TB, 2, 1>, 2 * TB, 2, 1);
            }
            if (B->tabqf_len)
                launch_nt(ctx, K_POST_SMOOTH, k_post_fused<VT, ZT, 4 * TB, 1, 1>, B->tabqf_len, 4 * TB,
                          fuse_lds(4 * TB, 1, 1, true), S, (const int2*)B->tabqf);
        } else {
            if (mx <= 64)
                go(k_post_fused<VT, ZT, TB, 1, 1>, TB, 1, 1);
            else if (mx <= 127)
                go(k_post_fused<VT, ZT, TB, 2, 1>, TB, 2, 1);
            else
                go(k_post_fused<VT, ZT, TB, 2, 2>, TB, 2, 2);
        }
    }
    if (nu > 0) {
        launch(ctx, K_POST_BLACK, k_post_black<C0, ZT, VT>, nu, B->lds0(), S, TU);
        launch(ctx, K_POST_RED, k_post_red<C0, ZT>, nu, B->lds0(), S, TU, 0);
        for (int v = 1; v < B->maxnu0_u; ++v) {
            launch(ctx, K_RB_BLACK, k_rb_black<C0, ZT>, nu, B->lds0(), S, TU, v);
            launch(ctx, K_POST_RED, k_post_red<C0, ZT>, nu, B->lds0(), S, TU, v);
        }
    }
}

void vcycle(phg_ctx* ctx, BatchDev* B, int first) {
    // the mixed precisions keep z in FP16 (scaled to an RMS near 1, so the
    // range of FP16 is ample and its 11 bits beat BF16's 8)
    if (B->prec == PHG_PREC_FP32) {
        if (B->binary) vcycle_t<Mask01, float, float, float>(ctx, B, first);
        else vcycle_t<float, float, float, float>(ctx, B, first);
    } else if (B->prec == PHG_PREC_MIXED_F16) {
        if (B->binary) vcycle_t<Mask01, _Float16, _Float16, _Float16>(ctx, B, first);
        else vcycle_t<_Float16, _Float16, _Float16, _Float16>(ctx, B, first);
    } else {
        if (B->binary) vcycle_t<Mask01, _Float16, bf16_t, _Float16>(ctx, B, first);
        else vcycle_t<_Float16, _Float16, bf16_t, _Float16>(ctx, B, first);
    }
}

// d = z + beta d_old, q = A d: on 2 TB-column blocks (the even column tiles)
// unless the batch asks for TB; one halo slot per thread when the rows allow.
// In the mixed precisions, 0/1 grids with even planes up to TB wide take the
// column-pair form by default (a batch asking for 2 TB threads keeps one
// column per thread).  FP32 keeps one column per thread: its 8-byte pairs
// measured 4% slower than 4-byte accesses, while FP16 pairs are 7% faster.
template <class OT, class ZT>
void dir_apply_t(phg_ctx* ctx, BatchDev* B) {
    const DSys* S = B->d_sys;
    if (std::is_same<OT, Mask01>::value && std::is_same<ZT, _Float16>::value && B->pair_dir) {
        const int nxe = (B->maxnx + 1) & ~1;
        const size_t lds = 2 * (size_t)(2 * TB + 2 * nxe) * sizeof(float);
        launch(ctx, K_DIR_APPLY, k_dir_apply_pair<ZT>, B->tabw_len, lds, S, (const int2*)B->tabw);
    } else if (B->fuse_nt != TB) {
        const size_t lds = 2 * (size_t)(2 * TB + 2 * B->maxnx) * sizeof(float);
        const int2* TW = B->tabw;
        if (2 * B->maxnx <= 2 * TB)
            launch_nt(ctx, K_DIR_APPLY, k_dir_apply<OT, ZT, 2 * TB, 1>, B->tabw_len, 2 * TB, lds, S, TW);
        else
            launch_nt(ctx, K_DIR_APPLY, k_dir_apply<OT, ZT, 2 * TB, MARCH_HALO>, B->tabw_len, 2 * TB, lds, S, TW);
    } else if (2 * B->maxnx <= TB) {
        launch(ctx, K_DIR_APPLY, k_dir_apply<OT, ZT, TB, 1>, B->tab_len[0], B->lds0(), S, B->tab[0]);
    } else {
        launch(ctx, K_DIR_APPLY, k_dir_apply<OT, ZT>, B->tab_len[0], B->lds0(), S, B->tab[0]);
    }
}

template <class OT>
void dir_apply(phg_ctx* ctx, BatchDev* B) {
    if (B->prec == PHG_PREC_FP32) dir_apply_t<OT, float>(ctx, B);
    else dir_apply_t<OT, _Float16>(ctx, B);
}

void cg_iteration(phg_ctx* ctx, BatchDev* B) {
    const DSys* S = B->d_sys;
    if (B->binary) dir_apply<Mask01>(ctx, B);
    else dir_apply<float>(ctx, B);
    launch(ctx, K_FIN, k_fin_alpha, B->nsys, 0, S, B->tab_sys);
    vcycle(ctx, B, 0);
    launch(ctx, K_FIN, k_fin_beta, B->nsys, 0, S, B->tab_sys);
    if (B->tab_z_len > 0)
        launch(ctx, K_ZONE, k_proj, (B->tab_z_len + PROJ_TASKS - 1) / PROJ_TASKS, 0, S, (const int*)B->state,
               (const int2*)B->tab_z, B->tab_z_len);
}

// closed-zone mean removal of every system's x (whatever its state)
// remove the closed-zone means of x (which = 0) or of the residual of the
// restarted systems (which = 1)
void demean_batch(phg_ctx* ctx, BatchDev* B, int which = 0) {
    if (B->tab_z_len == 0) return;
    const DSys* S = B->d_sys;
    launch(ctx, K_ZONE, k_zone_part, B->tab_z_len, 0, S, (const int2*)B->tab_z, which);
    launch(ctx, K_ZONE, k_zone_fin, B->nsys, 0, S, (const int2*)B->tab_sys, which);
    launch(ctx, K_ZONE, k_zone_sub, B->tab_len[0], 0, S, (const int2*)B->tab[0], which);
}

// r_0 = b - A x_0, the stop test at iteration 0, z_0 = M^-1 r_0, r.z; a
// restart projects its residual onto the range of A first.  hi: residuals
// formed from x accumulate in FP64.
void solve_setup(phg_ctx* ctx, BatchDev* B, bool restart, int hi) {
    const DSys* S = B->d_sys;
    if (B->binary) launch(ctx, K_RESID_INIT, k_resid_init<Mask01>, B->tab_len[0], 0, S, B->tab[0], hi);
    else launch(ctx, K_RESID_INIT, k_resid_init<float>, B->tab_len[0], 0, S, B->tab[0], hi);
    if (restart) demean_batch(ctx, B, 1);
    launch(ctx, K_FIN, k_fin_init, B->nsys, 0, S, B->tab_sys);
    vcycle(ctx, B, 1);
    launch(ctx, K_FIN, k_fin_rz0, B->nsys, 0, S, B->tab_sys);
}

int upload_sys(phg_ctx* ctx, BatchDev* B) {
    HIP_OK(hipMemcpyAsync(B->d_sys, B->h_sys.data(), sizeof(DSys) * B->nsys, hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

// set every system active with the given rule; scal/state from the host
int reset_state(phg_ctx* ctx, BatchDev* B, const phg_solve_params* P) {
    std::vector<double> scal((size_t)B->nsys * S_NSCAL, 0.0);
    std::vector<int> state((size_t)B->nsys * ST_NSTATE, 0);
    for (int s = 0; s < B->nsys; ++s) {
        double* sc = &scal[(size_t)s * S_NSCAL];
        int* st = &state[(size_t)s * ST_NSTATE];
        sc[S_TOL] = P ? P->tol : 0.0;
        sc[S_INV] = P ? P->inv : 1.0f;
        sc[S_RZ] = 1.0;
        sc[S_ZS] = sc[S_ZSN] = 1.0;
        sc[S_PW] = (double)VC_LEVEL_SCALE * (P && P->coarse_weight > 0 ? (double)P->coarse_weight : 1.0);
        st[ST_ACTIVE] = 1;
        st[ST_MAXIT] = P ? P->max_iters : 0;
        st[ST_RULE] = P && P->stop_rule == PHG_STOP_MAXNORM ? 1 : 0;
        st[ST_ZEROX] = P && P->zero_x ? 1 : 0;
        st[ST_FLEX] = B->prec != PHG_PREC_FP32 ? 1 : 0;
    }
    HIP_OK(hipMemcpyAsync(B->scal, scal.data(), scal.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(B->state, state.data(), state.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // pageable sources: wait so the host vectors may go out of scope
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return 0;
}

GridDev* get_grid(phg_ctx* ctx, int id) {
    auto it = ctx->grids.find(id);
    return it == ctx->grids.end() ? nullptr : it->second;
}
BatchDev* get_batch(phg_ctx* ctx, int id) {
    auto it = ctx->batches.find(id);
    return it == ctx->batches.end() ? nullptr : it->second;
}

// 0, or PHG_ESTATE when a grid of the batch was rebuilt by a face update
// after the batch was made (the batch holds the old set-up's buffers)
int batch_fresh(phg_ctx* ctx, const BatchDev* B, int batch) {
    for (int s = 0; s < B->nsys; ++s) {
        auto it = ctx->grids.find(B->grids[s]);
        if (it == ctx->grids.end())
            return fail(ctx, PHG_ESTATE, "batch %d: grid %d no longer exists", batch, B->grids[s]);
        if (it->second->gen != B->gens[s])
            return fail(ctx, PHG_ESTATE, "batch %d: grid %d was rebuilt by a face update; create the batch again",
                        batch, B->grids[s]);
    }
    return 0;
}

void destroy_batch(BatchDev* B) {
    if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->graph_full_exec) (void)hipGraphExecDestroy(B->graph_full_exec);
    if (B->graph_full) (void)hipGraphDestroy(B->graph_full);
    if (B->h_count) (void)hipHostFree(B->h_count);
    free_all(B->allocs);
    delete B;
}

void destroy_grid(GridDev* G) {
    free_all(G->crack_allocs);
    free_all(G->allocs);
    delete G;
}

int single_batch(phg_ctx* ctx, int grid, BatchDev** out) {
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (G->single_batch < 0) {
        int b = -1;
        int rc = phg_batch_create(ctx, 1, &grid, &b);
        if (rc) return rc;
        G->single_batch = b;
    }
    *out = get_batch(ctx, G->single_batch);
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int phg_create(phg_ctx** out, int device) {
    if (!out) return PHG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PHG_EHIP;
    if (device < 0 || device >= n) return PHG_EINVAL;
    phg_ctx* ctx = new phg_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PHG_EHIP;
    }
    // the device address of this device's workspace copy (device current)
    void* sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || sp == nullptr) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return PHG_EHIP;
    }
    ctx->scratch_dev = static_cast<double*>(sp);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        if (prop.sharedMemPerBlock > 0) ctx->lds_block = prop.sharedMemPerBlock;
        if (prop.multiProcessorCount > 0) ctx->cus = prop.multiProcessorCount;
    }
    *out = ctx;
    return PHG_OK;
}

void phg_destroy(phg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->batches) destroy_batch(kv.second);
    for (auto& kv : ctx->grids) destroy_grid(kv.second);
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* phg_last_error(phg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int phg_synchronize(phg_ctx* ctx) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (ctx->prof) prof_drain(ctx);
    return PHG_OK;
}

void* phg_stream(phg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int phg_malloc(phg_ctx* ctx, void** dptr, size_t bytes) {
    if (!ctx || !dptr) return PHG_EINVAL;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(ctx, PHG_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return PHG_OK;
}
int phg_free(phg_ctx* ctx, void* dptr) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipFree(dptr));
    return PHG_OK;
}
int phg_memcpy_h2d(phg_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return PHG_OK;
}
int phg_memcpy_d2h(phg_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return PHG_OK;
}
int phg_memset(phg_ctx* ctx, void* dst, int value, size_t bytes) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return PHG_OK;
}

// ---------------------------------------------------------------------------
}  // extern "C"

// the host half of a grid's set-up: hierarchy, dense coarsest inverse, zones
struct phg_grid_plan {
    int nx = 0, ny = 0, nz = 0;
    bool used = false;
    HostHierarchy H;
    std::vector<int> zone;
    std::vector<uint8_t> zone_open;
};

namespace {

// the checks both halves make; the message is empty when the shape is fine
std::string check_shape(int nx, int ny, int nz) {
    char buf[128];
    if (nx < 2 || ny < 2 || nz < 2 || (long)nx * ny * nz > (1L << 30)) {
        snprintf(buf, sizeof buf, "bad grid %dx%dx%d", nx, ny, nz);
        return buf;
    }
    if (nx > MARCH_MAX_NX) {
        snprintf(buf, sizeof buf, "nx=%d exceeds the limit %d", nx, MARCH_MAX_NX);
        return buf;
    }
    return "";
}

// build_hierarchy and compute_zones into P; false with a message on failure
bool plan_geometry(int nx, int ny, int nz, const float* ax, const float* ay, const float* az, const uint8_t* blocked,
                   const uint8_t* wind_closed, phg_grid_plan& P, std::string* err) {
    P.nx = nx;
    P.ny = ny;
    P.nz = nz;
    if (!build_hierarchy(nx, ny, nz, ax, ay, az, blocked, wind_closed, kDenseMax, P.H, err)) return false;
    compute_zones(nx, ny, nz, ax, ay, az, blocked, P.zone, P.zone_open);
    return true;
}

}  // namespace

extern "C" {

int phg_grid_plan_build(int nx, int ny, int nz, const float* ax, const float* ay, const float* az,
                        const uint8_t* blocked, const uint8_t* wind_closed, phg_grid_plan** plan_out) {
    return guarded(nullptr, [&]() -> int {
        if (!plan_out) return PHG_EINVAL;
        *plan_out = nullptr;
        if (!ax || !ay || !az || !blocked || !check_shape(nx, ny, nz).empty()) return PHG_EINVAL;
        phg_grid_plan* P = new phg_grid_plan();
        std::string err;
        if (!plan_geometry(nx, ny, nz, ax, ay, az, blocked, wind_closed, *P, &err)) {
            delete P;
            return PHG_EINVAL;
        }
        *plan_out = P;
        return PHG_OK;
    });
}

void phg_grid_plan_free(phg_grid_plan* plan) { delete plan; }

int phg_grid_create(phg_ctx* ctx, int nx, int ny, int nz, float dx, const float* fx, const float* fy, const float* fz,
                    const float* ax, const float* ay, const float* az, const uint8_t* blocked,
                    const uint8_t* wind_closed, int precision, int* grid_out) {
    return phg_grid_create_planned(ctx, nx, ny, nz, dx, fx, fy, fz, ax, ay, az, blocked, wind_closed, precision,
                                   nullptr, grid_out);
}

}  // extern "C"

namespace {

// A grid's device set-up from host geometry (the work of phg_grid_create),
// not yet registered with the context.
int build_grid_dev(phg_ctx* ctx, int nx, int ny, int nz, float dx, const float* fx, const float* fy, const float* fz,
                   const float* ax, const float* ay, const float* az, const uint8_t* blocked,
                   const uint8_t* wind_closed, int precision, phg_grid_plan* plan, GridDev** out) {
        *out = nullptr;
        const std::string bad = check_shape(nx, ny, nz);
        if (!bad.empty()) return fail(ctx, PHG_EINVAL, "%s", bad.c_str());
        if (!(dx > 0)) return fail(ctx, PHG_EINVAL, "bad cell size dx=%g", dx);
        if (!fx || !fy || !fz || !ax || !ay || !az || !blocked) return fail(ctx, PHG_EINVAL, "null geometry array");
        if (precision < PHG_PREC_FP32 || precision > PHG_PREC_MIXED_BF16)
            return fail(ctx, PHG_EINVAL, "unknown precision %d", precision);
        if (plan && (plan->used || plan->nx != nx || plan->ny != ny || plan->nz != nz))
            return fail(ctx, PHG_EINVAL, "the plan is used or was built for another grid shape");
        HIP_OK(hipSetDevice(ctx->device));
        GridDev* G = new GridDev();
        G->nx = nx;
        G->ny = ny;
        G->nz = nz;
        G->n = nx * ny * nz;
        G->dx = dx;
        G->prec = precision;
        if (plan) {
            G->H = std::move(plan->H);
            G->zone = std::move(plan->zone);
            G->zone_open = std::move(plan->zone_open);
            plan->used = true;
        } else {
            phg_grid_plan P;
            std::string err;
            if (!plan_geometry(nx, ny, nz, ax, ay, az, blocked, wind_closed, P, &err)) {
                delete G;
                return fail(ctx, PHG_EINVAL, "%s", err.c_str());
            }
            G->H = std::move(P.H);
            G->zone = std::move(P.zone);
            G->zone_open = std::move(P.zone_open);
        }
        const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
        const HostLevel& L0 = G->H.L[0];
        int rc = 0;
    #define UP(dst, src, cnt) \
        if (!rc) rc = upload(ctx, G->allocs, &dst, src, cnt);
        UP(G->ax, L0.ax.data(), nfx)
        UP(G->ay, L0.ay.data(), nfy)
        UP(G->az, L0.az.data(), nfz)
        UP(G->fx, fx, nfx)
        UP(G->fy, fy, nfy)
        UP(G->fz, fz, nfz)
        UP(G->blocked, blocked, (size_t)G->n)
        if (!rc) {
            if (precision == PHG_PREC_FP32) {
                G->axc = G->ax;
                G->ayc = G->ay;
                G->azc = G->az;
            } else {
                rc = upload_ct(ctx, G->allocs, &G->axc, L0.ax, precision);
                if (!rc) rc = upload_ct(ctx, G->allocs, &G->ayc, L0.ay, precision);
                if (!rc) rc = upload_ct(ctx, G->allocs, &G->azc, L0.az, precision);
            }
        }
        const int nl = (int)G->H.L.size();
        G->lv.resize(nl);
        for (int l = 0; l < nl && !rc; ++l) {
            const HostLevel& L = G->H.L[l];
            GridDev::Lev& D = G->lv[l];
            D.n = L.n;
            D.ellk = L.ellk;
            if (l > 0) {
                UP(D.col, L.ell_col.data(), L.ell_col.size())
                if (!rc) rc = upload_ct(ctx, G->allocs, &D.val, L.ell_val, precision);
                if (!rc) rc = upload_ct(ctx, G->allocs, &D.diag, L.diag, precision);
            }
            if (l + 1 < nl) {
                UP(D.agg, L.agg.data(), L.agg.size())
                UP(D.cptr, L.cptr.data(), L.cptr.size())
                UP(D.cidx, L.cidx.data(), L.cidx.size())
            }
            if (l > 0 && L.n < 65536 && !rc) {
                // one load per ELL entry: (column << 16 | FP16 value bits), or
                // (column, FP32 value bits) in the FP32 precision
                const size_t m = L.ell_col.size();
                std::vector<uint32_t> pk(precision == PHG_PREC_FP32 ? 2 * m : m);
                for (size_t i = 0; i < m; ++i) {
                    if (precision == PHG_PREC_FP32) {
                        pk[2 * i] = (uint32_t)L.ell_col[i];
                        std::memcpy(&pk[2 * i + 1], &L.ell_val[i], 4);
                    } else {
                        pk[i] = (uint32_t)L.ell_col[i] << 16 | f32_to_f16_bits(L.ell_val[i]);
                    }
                }
                uint32_t* q = nullptr;
                rc = upload(ctx, G->allocs, &q, pk.data(), pk.size());
                D.cv = q;
            }
            if (l == 0 && nl > 1 && G->n < (1 << 24) && nx % 2 == 0 && !rc) {
                // each level-1 aggregate lies in one 2x2x2 block of cells: its
                // corner cell | the mask of its children among the block's eight
                // (bit dx | dy << 1 | dz << 2; ascending bits are ascending cells);
                // for even nx only, where every corner cell is even (the kernels
                // read a block's x-pairs as aligned pairs)
                const int n1 = G->H.L[1].n, sxy = nx * ny;
                std::vector<uint32_t> cb(n1);
                bool ok = true;
                for (int I = 0; I < n1 && ok; ++I) {
                    int corner = -1, last = -1;
                    uint32_t m = 0;
                    for (int p = L.cptr[I]; p < L.cptr[I + 1] && ok; ++p) {
                        const int c = L.cidx[p], k = c / sxy, j = (c - k * sxy) / nx, i = c - k * sxy - j * nx;
                        const int cc = (i & ~1) + nx * ((j & ~1) + ny * (k & ~1));
                        const int bit = (i & 1) | (j & 1) << 1 | (k & 1) << 2;
                        if (corner < 0) corner = cc;
                        ok = cc == corner && bit > last;
                        last = bit;
                        m |= 1u << bit;
                    }
                    ok = ok && corner >= 0;
                    cb[I] = (uint32_t)corner | m << 24;
                }
                if (ok) UP(D.cblk, cb.data(), cb.size())
            }
        }
        G->dense = !G->H.cinv.empty();
        G->coarse = G->dense ? PHG_COARSE_EXACT : PHG_COARSE_GS20;
        if (!G->H.gs_ord.empty()) {
            UP(G->gs_ord, G->H.gs_ord.data(), G->H.gs_ord.size())
            UP(G->gs_lptr, G->H.gs_lptr.data(), G->H.gs_lptr.size())
        }
        if (!G->dense) {
            // the coarsest level is left to the Gauss-Seidel sweeps
        } else if (G->H.nc > CP_MIN) {
            // the upper triangle in tiles (I <= J) of 64x64, each column-major
            const int nc = G->H.nc, nt = (nc + 63) / 64;
            std::vector<double> tp((size_t)nt * (nt + 1) / 2 * 4096, 0.0);
            size_t tix = 0;
            for (int I = 0; I < nt; ++I)
                for (int J = I; J < nt; ++J, ++tix)
                    for (int c = 0; c < 64 && 64 * J + c < nc; ++c)
                        for (int r = 0; r < 64 && 64 * I + r < nc; ++r)
                            tp[tix * 4096 + (size_t)c * 64 + r] = G->H.cinv[(size_t)(64 * I + r) * nc + 64 * J + c];
            UP(G->cinvp, tp.data(), tp.size())
            G->cnt = nt;
        } else {
            UP(G->cinv, G->H.cinv.data(), G->H.cinv.size())
        }
        {
            std::vector<float> invd(G->n, 0.0f);
            for (int c = 0; c < G->n; ++c)
                if (L0.diag[c] > 0) invd[c] = 1.0f / L0.diag[c];
            UP(G->invd, invd.data(), invd.size())
            // face mask: exact when every level-0 coefficient is 0 or 1
            bool binary = true;
            for (const std::vector<float>* a : {&L0.ax, &L0.ay, &L0.az})
                for (float v : *a)
                    if (v != 0.0f && v != 1.0f) binary = false;
            if (binary) {
                std::vector<uint8_t> m(G->n, 0);
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            const int c = i + nx * (j + ny * k);
                            const int xf = i + (nx + 1) * (j + ny * k), yf = i + nx * (j + (ny + 1) * k);
                            const float a[6] = {L0.ax[xf], L0.ax[xf + 1], L0.ay[yf], L0.ay[yf + nx], L0.az[c],
                                                L0.az[c + nx * ny]};
                            uint8_t b = 0;
                            for (int d = 0; d < 6; ++d)
                                if (a[d] > 0) b |= (uint8_t)(1u << d);
                            m[c] = b;
                        }
                UP(G->mask, m.data(), m.size())
                G->binary = true;
                // the fused post-smoothing reads a cell's aggregate and mask
                // in one load: aggregates under 2^24 - 1 (0xFFFFFF: none)
                if (G->H.L.size() > 1 && G->H.L[1].n < 0xFFFFFF) {
                    const std::vector<int>& agg = G->H.L[0].agg;
                    std::vector<uint32_t> am(G->n);
                    for (int c = 0; c < G->n; ++c)
                        am[c] = (agg[c] < 0 ? 0xFFFFFFu : (uint32_t)agg[c]) | (uint32_t)m[c] << 24;
                    UP(G->aggm, am.data(), am.size())
                }
            }
        }
        // closed zones: cell lists, cut into chunks, for the deterministic mean removal
        {
            std::vector<int> zid(G->zone_open.size(), -1);
            int nclosed = 0;
            for (size_t z = 0; z < G->zone_open.size(); ++z)
                if (!G->zone_open[z]) zid[z] = nclosed++;
            std::vector<int> zptr(nclosed + 1, 0), zclosed(G->n, -1);
            for (int c = 0; c < G->n; ++c)
                if (G->zone[c] >= 0 && zid[G->zone[c]] >= 0) {
                    zclosed[c] = zid[G->zone[c]];
                    zptr[zclosed[c] + 1]++;
                }
            for (int z = 0; z < nclosed; ++z) zptr[z + 1] += zptr[z];
            std::vector<int> zcells(std::max(zptr[nclosed], 1)), fill(zptr.begin(), zptr.end() - 1);
            for (int c = 0; c < G->n; ++c)
                if (zclosed[c] >= 0) zcells[fill[zclosed[c]]++] = c;
            G->nclosed = nclosed;
            std::vector<int4> tasks;
            std::vector<int> ztptr(nclosed + 1, 0);
            for (int z = 0; z < nclosed; ++z) {
                ztptr[z] = (int)tasks.size();
                for (int b = zptr[z]; b < zptr[z + 1]; b += kZoneChunk)
                    tasks.push_back(make_int4(z, b, std::min(zptr[z + 1], b + kZoneChunk), 0));
            }
            ztptr[nclosed] = (int)tasks.size();
            G->nztask = (int)tasks.size();
            if (tasks.empty()) tasks.push_back(make_int4(0, 0, 0, 0));
            UP(G->ztask, tasks.data(), tasks.size())
            UP(G->ztptr, ztptr.data(), ztptr.size())
            UP(G->zcells, zcells.data(), zcells.size())
            UP(G->zclosed, zclosed.data(), zclosed.size())
            // volumes as the reference keeps them (cell count * dx^3 in FP64, gasphase.cpp:155-156)
            std::vector<int> czone(std::max(nclosed, 1), 0);
            std::vector<float> zvol(std::max(nclosed, 1), 0.0f);
            for (size_t z = 0; z < G->zone_open.size(); ++z)
                if (zid[z] >= 0) {
                    const double d = (double)dx;
                    czone[zid[z]] = (int)z;
                    zvol[zid[z]] = (float)((double)(zptr[zid[z] + 1] - zptr[zid[z]]) * d * d * d);
                }
            UP(G->czone, czone.data(), czone.size())
            UP(G->zvol, zvol.data(), zvol.size())
        }
        if (!rc) rc = dalloc(ctx, G->allocs, &G->scratch_b, (size_t)G->n);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->vmax, 4);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->lit_new, (size_t)G->n);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->cpart, (size_t)tiles(G->n));
        if (!rc) rc = dalloc(ctx, G->allocs, &G->zsum, (size_t)(G->nztask + G->nclosed + 1));
        if (!rc) rc = dalloc(ctx, G->allocs, &G->zval, (size_t)(G->nclosed + 1));
        if (!rc) rc = dalloc(ctx, G->allocs, &G->wind, (size_t)4 * nz);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->zp, G->zone_open.size() + 1);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->zpn, G->zone_open.size() + 1);
        if (!rc) rc = dalloc(ctx, G->allocs, &G->zdp, (size_t)G->nclosed + 1);
    #undef UP
        if (rc) {
            destroy_grid(G);
            return rc;
        }
        // what a later face update needs to redo the set-up: the wind-closed
        // sides and the caller's coefficients on them (level 0 holds zeros there)
        G->has_wc = wind_closed != nullptr;
        for (int q = 0; q < 4; ++q) {
            G->wc[q] = wind_closed ? wind_closed[q] : 0;
            G->side_a[q].clear();
            if (!G->wc[q]) continue;
            const int axis = q / 2, side = q % 2;
            if (axis == 0)
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j) G->side_a[q].push_back(ax[(side ? nx : 0) + (nx + 1) * (j + ny * k)]);
            else
                for (int k = 0; k < nz; ++k)
                    for (int i = 0; i < nx; ++i) G->side_a[q].push_back(ay[i + nx * ((side ? ny : 0) + (ny + 1) * k)]);
        }
        *out = G;
        return PHG_OK;
}

}  // namespace

extern "C" {

int phg_grid_create_planned(phg_ctx* ctx, int nx, int ny, int nz, float dx, const float* fx, const float* fy,
                            const float* fz, const float* ax, const float* ay, const float* az,
                            const uint8_t* blocked, const uint8_t* wind_closed, int precision, phg_grid_plan* plan,
                            int* grid_out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !grid_out) return PHG_EINVAL;
        GridDev* G = nullptr;
        const int rc = build_grid_dev(ctx, nx, ny, nz, dx, fx, fy, fz, ax, ay, az, blocked, wind_closed, precision,
                                      plan, &G);
        if (rc) return rc;
        const int id = ctx->next_id++;
        ctx->grids[id] = G;
        *grid_out = id;
        return PHG_OK;
    });
}

int phg_grid_set_coarse(phg_ctx* ctx, int grid, int mode) {
    if (!ctx) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (mode != PHG_COARSE_EXACT && mode != PHG_COARSE_GS20) return fail(ctx, PHG_EINVAL, "unknown coarse mode %d", mode);
    if (mode == PHG_COARSE_EXACT && !G->dense)
        return fail(ctx, PHG_EINVAL, "grid %d: the coarsest level has %d unknowns, too many for the dense inverse (%d)",
                    grid, G->H.nc, kDenseMax);
    if (mode == PHG_COARSE_GS20 && G->H.L.size() < 2)
        return fail(ctx, PHG_EINVAL, "grid %d has no coarse level (its dense solve is exact)", grid);
    if (mode != G->coarse && G->single_batch >= 0) {
        // the cached one-system batch was laid out for the other mode
        const int b = G->single_batch;
        int rc = phg_batch_destroy(ctx, b);
        if (rc) return rc;
    }
    G->coarse = mode;
    return PHG_OK;
}

int phg_grid_get_coarse(phg_ctx* ctx, int grid, int* mode) {
    if (!ctx || !mode) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    *mode = G->coarse;
    return PHG_OK;
}

int phg_grid_destroy(phg_ctx* ctx, int grid) {
    if (!ctx) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (G->single_batch >= 0) phg_batch_destroy(ctx, G->single_batch);
    destroy_grid(G);
    ctx->grids.erase(grid);
    return PHG_OK;
}

int phg_grid_info(phg_ctx* ctx, int grid, int* nlevels, int* level_n, int* nzones, int* ref_levels) {
    GridDev* G = ctx ? get_grid(ctx, grid) : nullptr;
    if (!G) return ctx ? fail(ctx, PHG_ESTATE, "unknown grid %d", grid) : PHG_EINVAL;
    if (nlevels) *nlevels = (int)G->H.L.size();
    if (level_n)
        for (size_t l = 0; l < G->H.L.size(); ++l) level_n[l] = G->H.L[l].n;
    if (nzones) *nzones = (int)G->zone_open.size();
    if (ref_levels) *ref_levels = G->H.ref_levels;
    return PHG_OK;
}

int phg_grid_zones(phg_ctx* ctx, int grid, int* zone, uint8_t* zone_open) {
    GridDev* G = ctx ? get_grid(ctx, grid) : nullptr;
    if (!G) return ctx ? fail(ctx, PHG_ESTATE, "unknown grid %d", grid) : PHG_EINVAL;
    if (zone) std::copy(G->zone.begin(), G->zone.end(), zone);
    if (zone_open) std::copy(G->zone_open.begin(), G->zone_open.end(), zone_open);
    return PHG_OK;
}

int phg_hierarchy_shape(int nx, int ny, int nz, const float* ax, const float* ay, const float* az,
                        const uint8_t* blocked, const uint8_t* wind_closed, int* nlevels, int* level_n, int* agg0) {
    return guarded(nullptr, [&]() -> int {
        if (!ax || !ay || !az || !blocked || !nlevels) return PHG_EINVAL;
        HostHierarchy H;
        std::string err;
        if (!build_hierarchy(nx, ny, nz, ax, ay, az, blocked, wind_closed, kDenseMax, H, &err)) return PHG_EINVAL;
        *nlevels = (int)H.L.size();
        if (level_n)
            for (size_t l = 0; l < H.L.size(); ++l) level_n[l] = H.L[l].n;
        if (agg0) std::copy(H.L[0].agg.begin(), H.L[0].agg.end(), agg0);
        return PHG_OK;
    });
}

// ---------------------------------------------------------------------------
int phg_batch_create(phg_ctx* ctx, int nsys, const int* grids, int* batch_out) {
    return phg_batch_create_ex(ctx, nsys, grids, nullptr, batch_out);
}

int phg_batch_create_ex(phg_ctx* ctx, int nsys, const int* grids, const phg_batch_opts* opts, int* batch_out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || nsys <= 0 || !grids || !batch_out) return PHG_EINVAL;
        if (opts && (opts->kchunk < 0 || opts->level0_blocks < 0 || opts->unfused < 0 || opts->unfused > 1 ||
                     opts->sweeps0 < 0 || opts->sweeps0 > kMaxSweeps0 ||
                     (opts->fused_threads != 0 && opts->fused_threads != TB && opts->fused_threads != 2 * TB &&
                      opts->fused_threads != 4 * TB) ||
                     opts->coarse_lds < 0 || opts->coarse_lds > 2))
            return fail(ctx, PHG_EINVAL, "bad batch options");
        if (opts)
            for (int s = 0; s < nsys; ++s) {
                if (opts->sys_kchunk && opts->sys_kchunk[s] < 0) return fail(ctx, PHG_EINVAL, "bad kchunk of system %d", s);
                if (opts->sys_sweeps0 && (opts->sys_sweeps0[s] < 0 || opts->sys_sweeps0[s] > kMaxSweeps0))
                    return fail(ctx, PHG_EINVAL, "bad sweep count of system %d", s);
            }
        HIP_OK(hipSetDevice(ctx->device));
        BatchDev* B = new BatchDev();
        B->nsys = nsys;
        B->grids.assign(grids, grids + nsys);
        B->gens.assign(nsys, 0);
        B->unfused = opts && opts->unfused;
        B->fuse_nt = opts ? opts->fused_threads : 0;
        B->h_sys.resize(nsys);
        int rc = 0;
        long col_tiles = 0;
        bool pair_even = true;
        for (int s = 0; s < nsys && !rc; ++s) {
            GridDev* G = get_grid(ctx, grids[s]);
            if (!G) {
                rc = fail(ctx, PHG_ESTATE, "unknown grid %d", grids[s]);
                break;
            }
            B->gens[s] = G->gen;
            if (s == 0) B->prec = G->prec;
            else if (G->prec != B->prec) {
                rc = fail(ctx, PHG_EINVAL, "all grids of a batch must share one precision");
                break;
            }
            B->maxl = std::max(B->maxl, (int)G->lv.size());
            if (G->coarse == PHG_COARSE_GS20) B->maxnc_gs = std::max(B->maxnc_gs, G->H.nc);
            else B->maxnc = std::max(B->maxnc, G->H.nc);
            B->binary = s == 0 ? G->binary : (B->binary && G->binary);
            B->maxnx = std::max(B->maxnx, G->nx);
            col_tiles += tiles(G->nx * G->ny);
            pair_even = pair_even && (G->nx * G->ny) % 2 == 0;
        }
        B->pair_dir = B->binary && pair_even && B->maxnx <= TB && B->fuse_nt == 0;
        if (rc) {
            delete B;
            return rc;
        }
        // planes per level-0 block: split the columns' k range until the launch
        // offers about kLevel0Blocks blocks
        const long want = opts && opts->level0_blocks > 0 ? opts->level0_blocks : kLevel0Blocks;
        const int nkc_target = (int)std::max(1L, (want + col_tiles - 1) / col_tiles);
        // a block's march pays a two-plane lead and its pipeline fill per chunk:
        // when the batch's columns alone give the chip enough blocks, chunks keep
        // at least kMinChunk planes (the 1.28M-cell house, 16 systems: 16-plane
        // chunks 772 solves/s, 40-plane 858)
        const bool cols_fill = col_tiles >= kColTilesFull && !(opts && opts->level0_blocks > 0);
        const size_t vts = vt_size(B->prec);
        if (!rc) rc = dalloc(ctx, B->allocs, &B->scal, (size_t)nsys * S_NSCAL);
        if (!rc) rc = dalloc(ctx, B->allocs, &B->state, (size_t)nsys * ST_NSTATE);
        if (!rc) rc = dalloc(ctx, B->allocs, &B->d_count, 1);
        if (!rc && hipHostMalloc((void**)&B->h_count, sizeof(int), hipHostMallocDefault) != hipSuccess)
            rc = fail(ctx, PHG_ENOMEM, "hipHostMalloc failed");
        for (int s = 0; s < nsys && !rc; ++s) {
            GridDev* G = get_grid(ctx, grids[s]);
            DSys& S = B->h_sys[s];
            std::memset(&S, 0, sizeof(DSys));
            S.nl = (int)G->lv.size();
            S.nc = G->H.nc;
            const int kc = opts && opts->sys_kchunk && opts->sys_kchunk[s] > 0 ? opts->sys_kchunk[s]
                           : opts && opts->kchunk > 0                        ? opts->kchunk
                                                                             : 0;
            const int nkc = cols_fill ? std::max(1, std::min(nkc_target, G->nz / kMinChunk)) : nkc_target;
            S.kchunk = kc > 0 ? std::min(kc, G->nz) : std::max(1, (G->nz + nkc - 1) / nkc);
            const int nu = opts && opts->sys_sweeps0 && opts->sys_sweeps0[s] > 0 ? opts->sys_sweeps0[s]
                           : opts && opts->sweeps0 > 0                          ? opts->sweeps0
                                                                                : 1;
            S.nu0 = nu;
            S.tiles0 = tiles(G->nx * G->ny) * ((G->nz + S.kchunk - 1) / S.kchunk);
            S.cinv = G->cinv;
            S.cinvp = G->cinvp;
            S.cnt = G->cnt;
            S.gs_ord = G->gs_ord;
            S.gs_lptr = G->gs_lptr;
            S.gs_nf = G->H.gs_nf;
            S.gs_nb = G->H.gs_nb;
            S.invd = G->invd;
            S.scal = B->scal + (size_t)s * S_NSCAL;
            S.state = B->state + (size_t)s * ST_NSTATE;
            S.nclosed = G->nclosed;
            S.nztask = G->nztask;
            S.ztask = G->ztask;
            S.ztptr = G->ztptr;
            S.zcells = G->zcells;
            S.zclosed = G->zclosed;
            const size_t n0 = G->n;
            float* v4 = nullptr;
            rc = dalloc(ctx, B->allocs, &v4, 7 * n0);
            if (rc) break;
            S.r = v4;
            S.q = v4 + n0;
            S.d0 = v4 + 2 * n0;
            S.d1 = v4 + 3 * n0;
            S.z = v4 + 4 * n0;   // the level-0 iterate: FP32, or FP16 in the mixed precisions
            S.r1 = v4 + 5 * n0;
            S.z2 = v4 + 6 * n0;
            rc = dalloc(ctx, B->allocs, &S.part, (size_t)S.tiles0 * NPART);
            if (!rc && S.cnt) rc = dalloc(ctx, B->allocs, &S.cpart, (size_t)S.cnt * S.cnt * 64);
            if (!rc) rc = dalloc(ctx, B->allocs, &S.zpart, (size_t)std::max(G->nztask, 1));
            if (!rc) rc = dalloc(ctx, B->allocs, &S.zmean, (size_t)std::max(G->nclosed, 1));
            if (rc) break;
            // V-cycle storage: level-0 residual scratch, coarse x b r t
            size_t total = n0;
            for (int l = 1; l < S.nl; ++l) total += 4 * (size_t)G->lv[l].n;
            char* slab = nullptr;
            rc = dalloc(ctx, B->allocs, &slab, total * vts);
            if (rc) break;
            for (int l = 0; l < S.nl; ++l) {
                DLevel& D = S.L[l];
                const GridDev::Lev& GL = G->lv[l];
                D.n = GL.n;
                D.ellk = GL.ellk;
                D.col = GL.col;
                D.val = GL.val;
                D.diag = GL.diag;
                D.agg = GL.agg;
                D.cptr = GL.cptr;
                D.cidx = GL.cidx;
                D.cv = GL.cv;
                D.cblk = GL.cblk;
                if (l == 0) {
                    D.nx = G->nx;
                    D.ny = G->ny;
                    D.nz = G->nz;
                    D.ax32 = G->ax;
                    D.ay32 = G->ay;
                    D.az32 = G->az;
                    D.axc = G->axc;
                    D.ayc = G->ayc;
                    D.azc = G->azc;
                    D.mask = G->mask;
                    D.aggm = G->aggm;
                    D.r = slab;
                    slab += n0 * vts;
                } else {
                    D.x = slab;
                    slab += GL.n * vts;
                    D.b = slab;
                    slab += GL.n * vts;
                    D.r = slab;
                    slab += GL.n * vts;
                    D.t = slab;
                    slab += GL.n * vts;
                }
            }
        }
        if (!rc) rc = dalloc(ctx, B->allocs, &B->d_sys, (size_t)nsys);
        // coarse levels in LDS: each system from its first level l (1 <= l <= nl-2)
        // whose down and up blocks fit the device's LDS per block; by default only
        // when at least two such blocks per CU would run.  (Smaller batches keep
        // the per-level launches: starting the LDS levels deeper, at a level of a
        // few thousand rows, measured 3-5% slower on 16- and 64-system batches.)
        B->tail_s.assign(nsys, 0);
        if (!rc) {
            const int mode = opts ? opts->coarse_lds : 0;
            auto first_level = [&](const GridDev* G) {
                const int nl = (int)G->lv.size();
                for (int l = 1; l + 2 <= nl; ++l) {
                    // every level below l but the coarsest must fit half the block's LDS
                    const bool fits = (l + 2 == nl || 2 * G->lv[l + 1].n <= G->lv[l].n) && G->lv[l].cv;
                    if (fits && tail_lds(G->lv[l].n, vts) <= ctx->lds_block) return l;
                }
                return nl;   // no LDS levels
            };
            long wide = 0;
            for (int s = 0; s < nsys; ++s) {
                const GridDev* G = get_grid(ctx, grids[s]);
                if (first_level(G) + 2 <= (int)G->lv.size()) ++wide;
            }
            const bool use = mode == 2 || (mode == 0 && wide >= 2L * ctx->cus);
            std::vector<int2> tt;
            for (int s = 0; s < nsys; ++s) {
                const GridDev* G = get_grid(ctx, grids[s]);
                const int nl = (int)G->lv.size();
                const int ts = use ? first_level(G) : nl;
                B->tail_s[s] = ts;
                if (ts + 2 <= nl) {
                    tt.push_back(make_int2(s, ts));
                    B->tail_lds = std::max(B->tail_lds, tail_lds(G->lv[ts].n, vts));
                }
            }
            B->tab_tail_len = (int)tt.size();
            if (!tt.empty()) rc = upload(ctx, B->allocs, &B->tab_tail, tt.data(), tt.size());
        }
        // tile tables: level 0 by column tiles x k-chunks; coarse levels by 256
        // unknowns of a group of up to CV_GRP consecutive systems on one grid
        B->tab.assign(1, nullptr);
        B->tab_len.assign(1, 0);
        B->tabg.assign(std::max(B->maxl, 1), nullptr);
        B->tabg_len.assign(std::max(B->maxl, 1), 0);
        if (!rc) {
            std::vector<int2> logical;
            for (int s = 0; s < nsys; ++s)
                for (int k = 0; k < B->h_sys[s].tiles0; ++k) logical.push_back(make_int2(s, k));
            std::vector<int2> phys = xcd_order(logical);
            B->tab_len[0] = (int)phys.size();
            rc = upload(ctx, B->allocs, &B->tab[0], phys.data(), phys.size());
            std::vector<int2> wide;
            for (const int2& e : logical) {
                const GridDev* G = get_grid(ctx, grids[e.x]);
                if ((e.y % tiles(G->nx * G->ny)) % 2 == 0) wide.push_back(e);
            }
            std::vector<int2> wphys = xcd_order(wide);
            B->tabw_len = (int)wphys.size();
            if (!rc) rc = upload(ctx, B->allocs, &B->tabw, wphys.data(), wphys.size());
            // the smoothing split: which systems take the fused passes
            std::vector<int2> lf, lwf, lqf, lu;
            for (const int2& e : logical) {
                const GridDev* G = get_grid(ctx, grids[e.x]);
                const bool f = B->binary && !B->unfused && G->nx <= FUSE_MAX_NX && (G->nx * G->ny) % 2 == 0 &&
                               B->h_sys[e.x].nu0 == 1 && G->aggm;
                if (f) {
                    lf.push_back(e);
                    const int h = e.y % tiles(G->nx * G->ny);
                    if (B->fuse_nt == 4 * TB || (B->fuse_nt == 0 && G->nx > 128)) {
                        if (h % 4 == 0) lqf.push_back(e);
                    } else {
                        if (h % 2 == 0) lwf.push_back(e);
                        B->maxnx_f = std::max(B->maxnx_f, G->nx);
                    }
                } else {
                    lu.push_back(e);
                    B->maxnu0_u = std::max(B->maxnu0_u, B->h_sys[e.x].nu0);
                }
            }
            std::vector<int2> pf = xcd_order(lf), pwf = xcd_order(lwf), pqf = xcd_order(lqf), pu = xcd_order(lu);
            B->tabf_len = (int)pf.size();
            B->tabwf_len = B->fuse_nt == TB ? 0 : (int)pwf.size();
            B->tabqf_len = (int)pqf.size();
            B->tabu_len = (int)pu.size();
            if (!rc && !pf.empty()) rc = upload(ctx, B->allocs, &B->tabf, pf.data(), pf.size());
            if (!rc && B->tabwf_len) rc = upload(ctx, B->allocs, &B->tabwf, pwf.data(), pwf.size());
            if (!rc && !pqf.empty()) rc = upload(ctx, B->allocs, &B->tabqf, pqf.data(), pqf.size());
            if (!rc && !pu.empty()) rc = upload(ctx, B->allocs, &B->tabu, pu.data(), pu.size());
        }
        for (int l = 1; l < B->maxl && !rc; ++l) {
            // groups of CV_GRP, or of 2 or 1 while the level's launches would offer
            // the chip fewer than kCoarseBlocks blocks (each system's arithmetic is
            // the same in any group); systems whose LDS levels cover level l (and
            // so the restriction into it) are left to k_vc_down / k_vc_up
            auto level_tiles = [&](int grp, std::vector<int4>* out) {
                long nb = 0;
                for (int s = 0; s < nsys;) {
                    GridDev* G = get_grid(ctx, grids[s]);
                    int cnt = 1;
                    while (s + cnt < nsys && cnt < grp && grids[s + cnt] == grids[s]) ++cnt;
                    if ((int)G->lv.size() > l && B->tail_s[s] > l) {
                        nb += tiles(G->lv[l].n);
                        if (out)
                            for (int k = 0; k < tiles(G->lv[l].n); ++k)
                                out->push_back(make_int4(s, cnt, k, (int)G->lv.size() | std::min(TB, G->lv[l].n - k * TB) << 8));
                    }
                    s += cnt;
                }
                return nb;
            };
            int grp = CV_GRP;
            while (grp > 1 && level_tiles(grp, nullptr) < kCoarseBlocks) grp /= 2;
            std::vector<int4> logical;
            level_tiles(grp, &logical);
            std::vector<int4> phys = xcd_order(logical);
            B->tabg_len[l] = (int)phys.size();
            if (!phys.empty()) rc = upload(ctx, B->allocs, &B->tabg[l], phys.data(), phys.size());
        }
        if (!rc) {
            std::vector<int2> ts(nsys), tz;
            for (int s = 0; s < nsys; ++s) {
                ts[s] = make_int2(s, 0);
                for (int k = 0; k < B->h_sys[s].nztask; ++k) tz.push_back(make_int2(s, k));
            }
            rc = upload(ctx, B->allocs, &B->tab_sys, ts.data(), ts.size());
            B->tab_z_len = (int)tz.size();
            if (!rc && !tz.empty()) rc = upload(ctx, B->allocs, &B->tab_z, tz.data(), tz.size());
        }
        if (!rc) {
            // coarsest-solve blocks: runs of up to cg_sys consecutive systems on one grid
            std::vector<int4> tc;
            // systems per block: the longest run on one grid rounded up to 1, 4, 8
            // or 16 (16 while their right-hand sides fit 64 KB of LDS, else 8), for
            // the grids whose inverse is read whole, then smaller while that launch
            // would offer the chip fewer than kCoarseBlocks blocks (the grouping
            // leaves every system's sums as they are); up to 8 for the packed ones
            // (systems whose grid takes the Gauss-Seidel sweeps are left out: one
            // block each, k_vc_coarse_gs)
            auto gs_of = [&](int s) { return get_grid(ctx, grids[s])->coarse == PHG_COARSE_GS20; };
            // a run of at least kMfmaMinSys systems on one grid whose inverse is
            // read whole takes the MFMA product (k_vc_coarse_mfma)
            auto mfma_run = [&](int s, int cnt) {
                return cnt >= kMfmaMinSys && !gs_of(s) && !get_grid(ctx, grids[s])->cnt;
            };
            int maxrun = 1, maxrun_p = 1;
            for (int s = 0; s < nsys;) {
                int cnt = 1;
                while (s + cnt < nsys && grids[s + cnt] == grids[s]) ++cnt;
                if (gs_of(s) || mfma_run(s, cnt)) {
                } else if (get_grid(ctx, grids[s])->cnt) {
                    maxrun_p = std::max(maxrun_p, cnt);
                } else {
                    maxrun = std::max(maxrun, cnt);
                }
                s += cnt;
            }
            const int cap = coarse_lds(CG_SYS_MAX, B->maxnc) <= 65536 ? CG_SYS_MAX : CG_SYS_MAX / 2;
            B->cg_sys = maxrun <= 1 ? 1 : maxrun <= 4 ? 4 : maxrun <= CG_SYS_MAX / 2 ? CG_SYS_MAX / 2 : cap;
            auto coarse_blocks = [&](int cgs) {
                long nb = 0;
                for (int s = 0; s < nsys;) {
                    int cnt = 1;
                    while (s + cnt < nsys && grids[s + cnt] == grids[s]) ++cnt;
                    const GridDev* G = get_grid(ctx, grids[s]);
                    if (!G->cnt && !gs_of(s) && !mfma_run(s, cnt))
                        nb += (long)((cnt + cgs - 1) / cgs) * ((G->H.nc + CG_ROWS - 1) / CG_ROWS);
                    s += cnt;
                }
                return nb;
            };
            while (B->cg_sys > 1 && coarse_blocks(B->cg_sys) < kCoarseBlocks)
                B->cg_sys = B->cg_sys > CG_SYS_MAX / 2 ? CG_SYS_MAX / 2 : B->cg_sys > 4 ? 4 : 1;
            B->cp_sys = maxrun_p <= 1 ? 1 : maxrun_p <= 4 ? 4 : 8;
            std::vector<int4> tp, tr, tm;
            std::vector<int2> tg;
            for (int s = 0; s < nsys;) {
                const GridDev* G = get_grid(ctx, grids[s]);
                int run = 1;
                while (s + run < nsys && grids[s + run] == grids[s]) ++run;
                if (mfma_run(s, run)) {
                    // row tile outer: after xcd_order each XCD reads its own
                    // eighth of the inverse for every system group
                    for (int t = 0; t < (G->H.nc + CM_ROWS - 1) / CM_ROWS; ++t)
                        for (int g0 = 0; g0 < run; g0 += 16 * kMfmaTiles)
                            tm.push_back(make_int4(s + g0, std::min(16 * kMfmaTiles, run - g0), t, 0));
                    s += run;
                    continue;
                }
                const int lim = G->cnt ? B->cp_sys : B->cg_sys;
                int cnt = 1;
                while (s + cnt < nsys && cnt < lim && grids[s + cnt] == grids[s]) ++cnt;
                if (gs_of(s)) {
                    for (int g = 0; g < cnt; ++g) tg.push_back(make_int2(s + g, 0));
                } else if (G->cnt) {
                    for (int I = 0; I < G->cnt; ++I) {
                        for (int J = I; J < G->cnt; ++J) tp.push_back(make_int4(s, cnt, I, J));
                        tr.push_back(make_int4(s, cnt, I, 0));
                    }
                } else {
                    for (int t = 0; t < (G->H.nc + CG_ROWS - 1) / CG_ROWS; ++t) tc.push_back(make_int4(s, cnt, t, 0));
                }
                s += cnt;
            }
            B->tab_coarse_len = (int)tc.size();
            B->tab_cm_len = (int)tm.size();
            tm = xcd_order(tm);
            if (!rc && !tm.empty()) rc = upload(ctx, B->allocs, &B->tab_cm, tm.data(), tm.size());
            B->tab_cp_len = (int)tp.size();
            B->tab_cr_len = (int)tr.size();
            B->tab_gs_len = (int)tg.size();
            if (!tc.empty()) rc = upload(ctx, B->allocs, &B->tab_coarse, tc.data(), tc.size());
            if (!rc && !tg.empty()) rc = upload(ctx, B->allocs, &B->tab_gs, tg.data(), tg.size());
            if (!rc && !tp.empty()) rc = upload(ctx, B->allocs, &B->tab_cp, tp.data(), tp.size());
            if (!rc && !tr.empty()) rc = upload(ctx, B->allocs, &B->tab_cr, tr.data(), tr.size());
        }
        if (!rc) rc = upload_sys(ctx, B);
        if (!rc) rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? 0 : fail(ctx, PHG_EHIP, "sync");
        if (rc) {
            destroy_batch(B);
            return rc;
        }
        const int id = ctx->next_id++;
        ctx->batches[id] = B;
        *batch_out = id;
        return PHG_OK;
    });
}

int phg_batch_destroy(phg_ctx* ctx, int batch) {
    if (!ctx) return PHG_EINVAL;
    BatchDev* B = get_batch(ctx, batch);
    if (!B) return fail(ctx, PHG_ESTATE, "unknown batch %d", batch);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    destroy_batch(B);
    ctx->batches.erase(batch);
    for (auto& kv : ctx->grids)
        if (kv.second->single_batch == batch) kv.second->single_batch = -1;
    return PHG_OK;
}

int phg_batch_solve(phg_ctx* ctx, int batch, const float* const* b, float* const* x, const phg_solve_params* P,
                    phg_solve_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !b || !x || !P) return PHG_EINVAL;
        BatchDev* B = get_batch(ctx, batch);
        if (!B) return fail(ctx, PHG_ESTATE, "unknown batch %d", batch);
        if (int rs = batch_fresh(ctx, B, batch)) return rs;
        if (!(P->tol >= 0) || P->max_iters < 0) return fail(ctx, PHG_EINVAL, "bad tolerance or iteration cap");
        if (P->stop_rule != PHG_STOP_REL_L2 && P->stop_rule != PHG_STOP_MAXNORM)
            return fail(ctx, PHG_EINVAL, "unknown stop rule %d", P->stop_rule);
        if (P->stop_rule == PHG_STOP_MAXNORM && !(P->inv > 0 && std::isfinite(P->inv)))
            return fail(ctx, PHG_EINVAL, "PHG_STOP_MAXNORM needs inv = dx^2/dt > 0 (got %g)", (double)P->inv);
        if (!(P->coarse_weight >= 0 && P->coarse_weight < 2))
            return fail(ctx, PHG_EINVAL, "coarse_weight must lie in [0, 2) (got %g)", (double)P->coarse_weight);
        HIP_OK(hipSetDevice(ctx->device));
        for (int s = 0; s < B->nsys; ++s) {
            if (!b[s] || !x[s]) return fail(ctx, PHG_EINVAL, "null vector for system %d", s);
            B->h_sys[s].b = b[s];
            B->h_sys[s].x = x[s];
        }
        int rc = upload_sys(ctx, B);
        if (!rc) rc = reset_state(ctx, B, P);
        if (rc) return rc;
        const DSys* S = B->d_sys;
        // with refinement on, every residual formed from x (a warm start's
        // first one as well as a restart's) accumulates in FP64: an FP32 one
        // is noise once it nears eps |A| |x|, where a warm start begins
        auto setup = [&](bool restart) { solve_setup(ctx, B, restart, restart || P->refine > 0 ? 1 : 0); };
        const int check = P->check_every > 0 ? P->check_every : 4;
        const bool graph = P->use_graph && !ctx->prof;
        const bool whole = P->use_graph == 2 && !ctx->prof;
        // after the whole-solve graph, the systems still running are the slow
        // tail: check a quarter as far apart (at least 4 iterations)
        const int tail_check = whole ? std::min(check, std::max(4, check / 4)) : check;
        // rounds: the first from x_0; each later one restarts the systems that
        // converged from their x and the true residual b - A x (those already
        // within tol stop at once); the cap applies per round
        for (int round = 0;; ++round) {
            int done = 0;
            bool live = true;
            if (whole && round == 0) {
                // the set-up and the first `check` iterations as one graph: one launch
                // per solve when every system stops within them (a system that has
                // stopped leaves each later kernel at once)
                const int N = std::min(check, P->max_iters);
                // the set-up's residual precision (refine) is fixed at capture
                const int hi0 = P->refine > 0 ? 1 : 0;
                if (!B->graph_full_exec || B->graph_full_n != N || B->graph_full_hi != hi0) {
                    if (B->graph_full_exec) (void)hipGraphExecDestroy(B->graph_full_exec);
                    if (B->graph_full) (void)hipGraphDestroy(B->graph_full);
                    B->graph_full_exec = nullptr;
                    B->graph_full = nullptr;
                    HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                    setup(false);
                    for (int i = 0; i < N; ++i) cg_iteration(ctx, B);
                    HIP_OK(hipStreamEndCapture(ctx->stream, &B->graph_full));
                    HIP_OK(hipGraphInstantiate(&B->graph_full_exec, B->graph_full, nullptr, nullptr, 0));
                    B->graph_full_n = N;
                    B->graph_full_hi = hi0;
                }
                HIP_OK(hipGraphLaunch(B->graph_full_exec, ctx->stream));
                done = N;
                launch(ctx, K_COUNT, k_count_active, 1, 0, S, B->nsys, B->d_count);
                HIP_OK(hipMemcpyAsync(B->h_count, B->d_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                HIP_OK(hipStreamSynchronize(ctx->stream));
                live = *B->h_count != 0;
            } else {
                setup(round > 0);
                if (round > 0) {   // a restart round: iterate only if some true residual missed tol
                    launch(ctx, K_COUNT, k_count_active, 1, 0, S, B->nsys, B->d_count);
                    HIP_OK(hipMemcpyAsync(B->h_count, B->d_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                    HIP_OK(hipStreamSynchronize(ctx->stream));
                    live = *B->h_count != 0;
                }
            }
            if (graph && live && !B->graph_exec) {
                HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                cg_iteration(ctx, B);
                HIP_OK(hipStreamEndCapture(ctx->stream, &B->graph));
                HIP_OK(hipGraphInstantiate(&B->graph_exec, B->graph, nullptr, nullptr, 0));
            }
            while (live && done < P->max_iters) {
                const int k = std::min(round == 0 ? tail_check : check, P->max_iters - done);
                for (int i = 0; i < k; ++i) {
                    if (graph) HIP_OK(hipGraphLaunch(B->graph_exec, ctx->stream));
                    else cg_iteration(ctx, B);
                }
                done += k;
                launch(ctx, K_COUNT, k_count_active, 1, 0, S, B->nsys, B->d_count);
                HIP_OK(hipMemcpyAsync(B->h_count, B->d_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                HIP_OK(hipStreamSynchronize(ctx->stream));
                if (*B->h_count == 0) break;
            }
            if (B->prec == PHG_PREC_FP32) launch(ctx, K_XFLUSH, k_x_flush<float>, B->tab_len[0], 0, S, B->tab[0]);
            else launch(ctx, K_XFLUSH, k_x_flush<_Float16>, B->tab_len[0], 0, S, B->tab[0]);
            if (round >= P->refine) break;
            launch(ctx, K_FIN, k_refine_reset, (B->nsys + TB - 1) / TB, 0, S, B->nsys);
        }
        if (P->demean) demean_batch(ctx, B);
        HIP_OK(hipGetLastError());
        if (stats) {
            std::vector<double> scal((size_t)B->nsys * S_NSCAL);
            std::vector<int> state((size_t)B->nsys * ST_NSTATE);
            HIP_OK(hipMemcpyAsync(scal.data(), B->scal, scal.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            HIP_OK(hipMemcpyAsync(state.data(), B->state, state.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_OK(hipStreamSynchronize(ctx->stream));
            for (int s = 0; s < B->nsys; ++s) {
                const double* sc = &scal[(size_t)s * S_NSCAL];
                const int* st = &state[(size_t)s * ST_NSTATE];
                stats[s].iters = st[ST_ITOT] + st[ST_ITERS];
                stats[s].status = st[ST_STATUS];
                stats[s].relres = sc[S_BB] > 0 ? std::sqrt(sc[S_RR]) / std::sqrt(sc[S_BB]) : 0.0;
                stats[s].maxres = (double)((float)sc[S_MAXR] / (float)sc[S_INV]);
            }
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
        if (ctx->prof) prof_drain(ctx);
        return PHG_OK;
    });
}

int phg_batch_graph_check(phg_ctx* ctx, int batch, int iterations, int refine, int* nodes_out, char* msg,
                          size_t msg_len) {
    if (!ctx || iterations < 0) return PHG_EINVAL;
    BatchDev* B = get_batch(ctx, batch);
    if (!B) return fail(ctx, PHG_ESTATE, "unknown batch %d", batch);
    if (int rs = batch_fresh(ctx, B, batch)) return rs;
    if (ctx->prof) return fail(ctx, PHG_EINVAL, "profiling is on: launches are not captured then");
    HIP_OK(hipSetDevice(ctx->device));
    // the whole-solve graph phg_batch_solve captures for use_graph = 2
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    solve_setup(ctx, B, false, refine > 0 ? 1 : 0);
    for (int i = 0; i < iterations; ++i) cg_iteration(ctx, B);
    HIP_OK(hipStreamEndCapture(ctx->stream, &g));
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    std::vector<hipGraphNode_t> nodes(n);
    if (e == hipSuccess && n) e = hipGraphGetNodes(g, nodes.data(), &n);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, ctx->device);
    std::string bad;
    size_t kernels = 0, max_lds = 0, max_threads = 0;
    for (size_t k = 0; e == hipSuccess && k < n && bad.empty(); ++k) {
        hipGraphNodeType t;
        e = hipGraphNodeGetType(nodes[k], &t);
        if (e != hipSuccess) break;
        if (t != hipGraphNodeTypeKernel) {
            bad = "node " + std::to_string(k) + " is not a kernel node (type " + std::to_string((int)t) + ")";
            break;
        }
        hipKernelNodeParams kp;
        e = hipGraphKernelNodeGetParams(nodes[k], &kp);
        if (e != hipSuccess) break;
        hipFuncAttributes fa;
        e = hipFuncGetAttributes(&fa, kp.func);
        if (e != hipSuccess) break;
        ++kernels;
        const size_t threads = (size_t)kp.blockDim.x * kp.blockDim.y * kp.blockDim.z;
        const size_t lds = fa.sharedSizeBytes + kp.sharedMemBytes;
        const unsigned long long work = (unsigned long long)kp.gridDim.x * kp.blockDim.x;
        max_lds = std::max(max_lds, lds);
        max_threads = std::max(max_threads, threads);
        char buf[256];
        if (threads == 0 || kp.gridDim.x == 0 || kp.gridDim.y == 0 || kp.gridDim.z == 0)
            snprintf(buf, sizeof buf, "node %zu: empty launch (grid %u, block %zu)", k, kp.gridDim.x, threads);
        else if ((int)threads > fa.maxThreadsPerBlock || (int)threads > prop.maxThreadsPerBlock)
            snprintf(buf, sizeof buf, "node %zu: %zu threads per block, limit %d", k, threads,
                     std::min(fa.maxThreadsPerBlock, prop.maxThreadsPerBlock));
        else if (lds > (size_t)prop.maxSharedMemoryPerMultiProcessor)
            snprintf(buf, sizeof buf, "node %zu: %zu B of LDS, limit %zu", k, lds,
                     (size_t)prop.maxSharedMemoryPerMultiProcessor);
        else if (work > 0xFFFFFFFFull)
            snprintf(buf, sizeof buf, "node %zu: grid of %llu work-items exceeds 2^32", k, work);
        else if (fa.localSizeBytes > 0)
            snprintf(buf, sizeof buf, "node %zu: %zu B of scratch per lane", k, fa.localSizeBytes);
        else
            buf[0] = 0;
        bad = buf;
    }
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(ctx, PHG_EHIP, "graph inspection: %s", hipGetErrorString(e));
    if (nodes_out) *nodes_out = (int)n;
    if (msg && msg_len) {
        snprintf(msg, msg_len, "%zu kernel nodes of %zu; largest block %zu threads, largest LDS %zu B (device: %d "
                 "threads, %zu B per CU)%s%s",
                 kernels, n, max_threads, max_lds, prop.maxThreadsPerBlock, (size_t)prop.maxSharedMemoryPerMultiProcessor,
                 bad.empty() ? "" : "; ", bad.c_str());
    }
    if (!bad.empty()) return fail(ctx, PHG_EINVAL, "%s", bad.c_str());
    return PHG_OK;
}

int phg_solve(phg_ctx* ctx, int grid, const float* b, float* x, const phg_solve_params* P, phg_solve_stats* stats) {
    if (!ctx) return PHG_EINVAL;
    BatchDev* B = nullptr;
    int rc = single_batch(ctx, grid, &B);
    if (rc) return rc;
    return phg_batch_solve(ctx, get_grid(ctx, grid)->single_batch, &b, &x, P, stats);
}

int phg_apply(phg_ctx* ctx, int grid, const float* x, float* y, double* dot) {
    if (!ctx || !x || !y) return PHG_EINVAL;
    BatchDev* B = nullptr;
    int rc = single_batch(ctx, grid, &B);
    if (rc) return rc;
    if (B->binary)
        launch(ctx, K_DIR_APPLY, k_apply_plain<Mask01>, B->tab_len[0], 0, (const DSys*)B->d_sys, (const int2*)B->tab[0], x, y);
    else
        launch(ctx, K_DIR_APPLY, k_apply_plain<float>, B->tab_len[0], 0, (const DSys*)B->d_sys, (const int2*)B->tab[0], x, y);
    launch(ctx, K_FIN, k_fin_dot, 1, 0, (const DSys*)B->d_sys, (const int2*)B->tab_sys);
    double sc[S_NSCAL];
    HIP_OK(hipMemcpyAsync(sc, B->scal, sizeof sc, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipGetLastError());
    if (dot) *dot = sc[S_DQ];
    return PHG_OK;
}

int phg_precondition(phg_ctx* ctx, int grid, const float* r, float* z) {
    if (!ctx || !r || !z) return PHG_EINVAL;
    BatchDev* B = nullptr;
    int rc = single_batch(ctx, grid, &B);
    if (rc) return rc;
    // the first V-cycle reads r_0 from the system's own residual buffer (the
    // level-0 kernels rely on its alignment)
    HIP_OK(hipMemcpyAsync(B->h_sys[0].r, r, get_grid(ctx, grid)->n * sizeof(float), hipMemcpyDeviceToDevice,
                          ctx->stream));
    phg_solve_params P{};
    rc = reset_state(ctx, B, &P);
    if (rc) return rc;
    vcycle(ctx, B, 1);
    const DSys* S = B->d_sys;
    if (B->prec == PHG_PREC_FP32) launch(ctx, K_POST_RED, k_z_out<float>, B->tab_len[0], 0, S, B->tab[0], z);
    else launch(ctx, K_POST_RED, k_z_out<_Float16>, B->tab_len[0], 0, S, B->tab[0], z);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipGetLastError());
    if (ctx->prof) prof_drain(ctx);
    return PHG_OK;
}

// ---------------------------------------------------------------------------
int phg_rhs(phg_ctx* ctx, int grid, float dt, const float* u, const float* v, const float* w, const float* S, float* b,
            float* diag, float* p) {
    if (!ctx || !u || !v || !w || !S || !b || !p) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (!(dt > 0)) return fail(ctx, PHG_EINVAL, "dt must be positive");
    const float inv = G->dx * G->dx / dt;
    launch(ctx, K_RHS, k_rhs, tiles(G->n), 0, G->dg(), G->dx, inv, u, v, w, S, b, diag, p);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_leakage(phg_ctx* ctx, int grid, float* S, float* shift) {
    if (!ctx || !S) return PHG_EINVAL;
    BatchDev* B = nullptr;
    int rc = single_batch(ctx, grid, &B);
    if (rc) return rc;
    GridDev* G = get_grid(ctx, grid);
    float* saved = B->h_sys[0].x;
    B->h_sys[0].x = S;   // the closed-zone mean removal of x, applied to S
    rc = upload_sys(ctx, B);
    B->h_sys[0].x = saved;
    if (rc) return rc;
    demean_batch(ctx, B);
    if (shift) {
        std::vector<double> zm(std::max(G->nclosed, 1));
        if (G->nclosed)
            HIP_OK(hipMemcpyAsync(zm.data(), B->h_sys[0].zmean, G->nclosed * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        int k = 0;
        for (size_t z = 0; z < G->zone_open.size(); ++z) shift[z] = G->zone_open[z] ? 0.0f : (float)zm[k++];
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_zone_mean(phg_ctx* ctx, int grid, float* p) {
    if (!ctx || !p) return PHG_EINVAL;
    BatchDev* B = nullptr;
    int rc = single_batch(ctx, grid, &B);
    if (rc) return rc;
    float* saved = B->h_sys[0].x;
    B->h_sys[0].x = p;
    rc = upload_sys(ctx, B);
    B->h_sys[0].x = saved;
    if (rc) return rc;
    demean_batch(ctx, B);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_zone_mean_noalloc(phg_ctx* ctx, int grid, float* p) {
    if (!ctx || !p) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const long need = (long)G->nztask + G->nclosed;
    if (need > kNoAllocDoubles)
        return fail(ctx, PHG_EINVAL,
                    "phg_zone_mean_noalloc: grid %d needs %ld workspace entries (%d tasks + %d closed zones), "
                    "limit %d",
                    grid, need, G->nztask, G->nclosed, kNoAllocDoubles);
    if (G->nclosed == 0) return PHG_OK;
    // partials first, then the means; every entry read is written earlier in
    // the same call, so nothing needs clearing
    double* ws = ctx->scratch_dev;
    hipLaunchKernelGGL(k_zone_part_ws, dim3(G->nztask), dim3(TB), 0, ctx->stream, (const int4*)G->ztask,
                       (const int*)G->zcells, (const float*)p, ws);
    hipLaunchKernelGGL(k_zone_fin_ws, dim3(1), dim3(TB), 0, ctx->stream, G->nclosed, G->nztask,
                       (const int*)G->ztptr, (const int4*)G->ztask, ws);
    hipLaunchKernelGGL(k_zone_sub_ws, dim3(tiles(G->n)), dim3(TB), 0, ctx->stream, G->n, (const int*)G->zclosed,
                       (const double*)(ws + G->nztask), p);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_correct(phg_ctx* ctx, int grid, float dt, const float* p, float* u, float* v, float* w, float* max_vel) {
    if (!ctx || !p || !u || !v || !w) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const int nfaces = (G->nx + 1) * G->ny * G->nz + G->nx * (G->ny + 1) * G->nz + G->n;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, G->vmax, 1, 0.0f);
    launch(ctx, K_CORRECT, k_correct, tiles(nfaces), 0, G->dg(), dt / G->dx, p, u, v, w, G->vmax);
    if (max_vel) {
        HIP_OK(hipMemcpyAsync(max_vel, G->vmax, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_project(phg_ctx* ctx, int grid, float dt, float* u, float* v, float* w, float* p, const float* S,
                const phg_solve_params* params, phg_solve_stats* stats, float* max_vel) {
    if (!ctx || !params) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    int rc = phg_rhs(ctx, grid, dt, u, v, w, S, G->scratch_b, nullptr, p);
    if (rc) return rc;
    phg_solve_params P = *params;
    P.demean = 1;
    P.inv = G->dx * G->dx / dt;
    rc = phg_solve(ctx, grid, G->scratch_b, p, &P, stats);
    if (rc) return rc;
    return phg_correct(ctx, grid, dt, p, u, v, w, max_vel);
}

int phg_energy_expansion(phg_ctx* ctx, int grid, float dt, float* T, float* S, const float* Pb, const float* Q,
                         const float* massSrc, float* max_div) {
    if (!ctx || !T || !S || !Pb || !Q || !massSrc) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, G->vmax + 2, 1, 0.0f);
    launch(ctx, K_ENERGY, k_energy, tiles(G->n), 0, G->n, G->dx, dt, (const uint8_t*)G->blocked, Pb, Q, massSrc, T, S,
           G->vmax + 2);
    if (max_div) {
        HIP_OK(hipMemcpyAsync(max_div, G->vmax + 2, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_advect_scalars(phg_ctx* ctx, int grid, float dt, const float* u, const float* v, const float* w,
                       const float* Pb, const float* nu, const float* const* fields_in, float* const* fields_out) {
    if (!ctx || !u || !v || !w || !Pb || !nu || !fields_in || !fields_out) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    TransportArgs A;
    for (int q = 0; q < 10; ++q) {
        if (!fields_in[q] || !fields_out[q]) return fail(ctx, PHG_EINVAL, "null field %d", q);
        for (int r = 0; r < 10; ++r)
            if ((const void*)fields_out[q] == (const void*)fields_in[r])
                return fail(ctx, PHG_EINVAL, "output field %d aliases input field %d", q, r);
        A.in[q] = fields_in[q];
        A.out[q] = fields_out[q];
    }
    A.nx = G->nx;
    A.ny = G->ny;
    A.nz = G->nz;
    A.dx = G->dx;
    A.dt = dt;
    A.u = u;
    A.v = v;
    A.w = w;
    A.fx = G->fx;
    A.fy = G->fy;
    A.fz = G->fz;
    A.blocked = G->blocked;
    A.Pb = Pb;
    A.nu = nu;
    // column tiles x k-chunks of at least PHG_TR_KMIN planes, about PHG_TR_BLOCKS blocks
    A.tiles_x = (G->nx + TR_TX - 1) / TR_TX;
    A.tiles_y = (G->ny + TR_TY - 1) / TR_TY;
    const int ntile = A.tiles_x * A.tiles_y;
    const int want = std::max(1, (PHG_TR_BLOCKS + ntile - 1) / ntile);
    A.kchunk = std::max(PHG_TR_KMIN, (G->nz + want - 1) / want);
    const int nchunk = (G->nz + A.kchunk - 1) / A.kchunk;
    launch_nt(ctx, K_ADVECT, k_transport, ntile * nchunk, TR_THREADS, 0, A);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

// ---------------------------------------------------------------------------
// the rest of the gas step
// ---------------------------------------------------------------------------
int phg_stable_dt(float dt_max, float profile_dt_max, float cfl, float dx, float max_vel, float max_nu, float max_div,
                  float* dt_out) {
    if (!dt_out) return PHG_EINVAL;
    // stableDt (gasphase.cpp:63-70), on the host with the reference's float arithmetic
    const float PR_T = 0.5f;
    float dt = std::min(dt_max, profile_dt_max);
    dt = std::min(dt, cfl * dx / std::max(max_vel, 0.1f));
    const float alpha = max_nu * (1.0f + 1.0f / PR_T);
    if (alpha > 0) dt = std::min(dt, 0.9f * dx * dx / (6.0f * alpha));
    if (max_div > 0) dt = std::min(dt, 0.25f / max_div);
    *dt_out = std::max(dt, 1e-3f);
    return PHG_OK;
}

int phg_turbulence(phg_ctx* ctx, int grid, const float* u, const float* v, const float* w, const float* T, float* nu,
                   float* max_nu) {
    if (!ctx || !u || !v || !w || !T || !nu) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    TurbArgs A;
    A.nx = G->nx;
    A.ny = G->ny;
    A.nz = G->nz;
    A.dx = G->dx;
    A.u = u;
    A.v = v;
    A.w = w;
    A.T = T;
    A.blocked = G->blocked;
    A.nu = nu;
    A.max_nu = G->vmax + 1;
    A.tiles_x = (G->nx + TU_TX - 1) / TU_TX;
    A.tiles_y = (G->ny + TU_TY - 1) / TU_TY;
    const int ntile = A.tiles_x * A.tiles_y;
    // about eight resident blocks per CU, chunks down to one plane (a block's
    // planes are a serial chain)
    const int want = std::max(1, (2048 + ntile - 1) / ntile);
    A.kchunk = std::max(1, (G->nz + want - 1) / want);
    const int nchunk = (G->nz + A.kchunk - 1) / A.kchunk;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, G->vmax + 1, 1, 0.0f);
    launch_nt(ctx, K_TURB, k_turbulence, ntile * nchunk, TU_THREADS, 0, A);
    if (max_nu) {
        HIP_OK(hipMemcpyAsync(max_nu, G->vmax + 1, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_heat_diffusion(phg_ctx* ctx, int grid, float dt, const float* T, const float* nu, const float* Pb, float* Q) {
    if (!ctx || !T || !nu || !Pb || !Q) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const CellGrid cg{G->nx, G->ny, G->nz, G->dx, G->fx, G->fy, G->fz, G->blocked};
    launch(ctx, K_HEAT, k_heat_diffusion, tiles(G->n), 0, cg, dt, T, nu, Pb, Q);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

namespace {
// windAt (gasphase.cpp:52-61) for every boundary level, on the host with the
// host C library, so the device applies the reference's own values
int wind_table(phg_ctx* ctx, GridDev* G, const phg_wind* wd) {
    const float key[4] = {wd->speed, wd->from_deg, wd->exponent, wd->oz};
    if (G->wind_set && std::memcmp(key, G->wind_key, sizeof key) == 0) return 0;
    std::vector<float> t((size_t)4 * G->nz, std::numeric_limits<float>::quiet_NaN());
    if (wd->speed > 0) {
        const float th = wd->from_deg * 3.14159265f / 180.0f;
        for (int axis = 0; axis < 2; ++axis)
            for (int k = 0; k < G->nz; ++k) {
                const float z = wd->oz + (k + 0.5f) * G->dx;
                const float U = wd->speed * std::pow(std::max(z, 0.05f) / 10.0f, wd->exponent);
                const float comp = axis == 0 ? -std::sin(th) * U : -std::cos(th) * U;
                if (comp > 0) t[(size_t)(axis * 2 + 0) * G->nz + k] = comp;   // enters through the low side
                if (comp < 0) t[(size_t)(axis * 2 + 1) * G->nz + k] = comp;   // through the high side
            }
    }
    HIP_OK(hipMemcpy(G->wind, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    std::memcpy(G->wind_key, key, sizeof key);
    G->wind_set = true;
    return 0;
}
}  // namespace

int phg_momentum(phg_ctx* ctx, int grid, float dt, const phg_wind* wind, const float* u, const float* v,
                 const float* w, const float* T, const float* Pb, const float* nu, const float* p, float* u_out,
                 float* v_out, float* w_out) {
    if (!ctx || !u || !v || !w || !T || !Pb || !nu || !p || !u_out || !v_out || !w_out) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (u_out == u || v_out == v || w_out == w) return fail(ctx, PHG_EINVAL, "momentum output aliases its input");
    const bool windy = wind && wind->speed > 0;
    if (windy) {
        int rc = wind_table(ctx, G, wind);
        if (rc) return rc;
    }
    FlowArgs A;
    A.nx = G->nx;
    A.ny = G->ny;
    A.nz = G->nz;
    A.dx = G->dx;
    A.dt = dt;
    A.u = u;
    A.v = v;
    A.w = w;
    A.fx = G->fx;
    A.fy = G->fy;
    A.fz = G->fz;
    A.lx = G->lx;
    A.ly = G->ly;
    A.lz = G->lz;
    A.T = T;
    A.Pb = Pb;
    A.nu = nu;
    A.p = p;
    A.wind = windy ? G->wind : nullptr;
    A.u2 = u_out;
    A.v2 = v_out;
    A.w2 = w_out;
    const int nx = G->nx, ny = G->ny, nz = G->nz;
    launch(ctx, K_MOMENTUM, k_momentum<0>, tiles((nx + 1) * ny * nz), 0, A);
    launch(ctx, K_MOMENTUM, k_momentum<1>, tiles(nx * (ny + 1) * nz), 0, A);
    launch(ctx, K_MOMENTUM, k_momentum<2>, tiles(nx * ny * (nz + 1)), 0, A);
    launch(ctx, K_MOMENTUM, k_momentum_edge<0>, tiles(2 * ny * nz), 0, A);
    launch(ctx, K_MOMENTUM, k_momentum_edge<1>, tiles(2 * nz * nx), 0, A);
    launch(ctx, K_MOMENTUM, k_momentum_edge<2>, tiles(2 * nx * ny), 0, A);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

namespace {
bool gas_ok(const phg_gas* g) {
    return g && g->T && g->Yo && g->Yf && g->Ys && g->Ya && g->Yp && g->Yc && g->Yw && g->Yco && g->Yai && g->Pb &&
           g->massSrc && g->Q && g->Qr && g->hrr && g->pilot && g->lit;
}
GasFields gas_fields(const phg_gas* g) {
    return GasFields{g->T,  g->Yo, g->Yf,      g->Ys, g->Ya, g->Yp,  g->Yc,    g->Yw,  g->Yco,
                     g->Yai, g->Pb, g->massSrc, g->Q,  g->Qr, g->hrr, g->pilot, g->lit};
}
}  // namespace

int phg_step_clear(phg_ctx* ctx, int grid, const phg_gas* gas) {
    if (!ctx || !gas_ok(gas)) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    launch(ctx, K_GASMISC, k_step_clear, tiles(G->n), 0, G->n, gas->Q, gas->Qr, gas->pilot, gas->massSrc);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_add_gas(phg_ctx* ctx, int grid, int ncells, const int* cells, float fuel_gen, float aerosol, float soot_pot,
                float ait, int set_pilot, const phg_gas* gas) {
    if (!ctx || ncells < 0 || (ncells > 0 && !cells) || !gas_ok(gas)) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (ncells == 0) return PHG_OK;
    launch(ctx, K_GASMISC, k_add_gas, (ncells + TB - 1) / TB, 0, ncells, cells, (const uint8_t*)G->blocked, G->dx,
           fuel_gen, aerosol, soot_pot, ait, set_pilot, gas_fields(gas));
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_combustion(phg_ctx* ctx, int grid, float dt, int radiation_on, const phg_gas* gas, double* hrr_total) {
    if (!ctx || !gas_ok(gas)) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (!(dt > 0)) return fail(ctx, PHG_EINVAL, "dt must be positive");
    // mixing-limited fraction and radiant share of the flame heat (solver.cpp:287-289)
    const float frac = 1.0f - std::exp(-dt / 0.1f);
    const float chiR = radiation_on ? 0.35f : 0.0f;
    const int nb = tiles(G->n);
    launch(ctx, K_COMBUST, k_combustion, nb, 0, G->nx, G->ny, G->nz, G->dx, dt, frac, chiR,
           (const uint8_t*)G->blocked, gas_fields(gas), G->lit_new, G->cpart);
    launch_nt(ctx, K_COMBUST, k_copy_u8, (G->n + TB - 1) / TB, TB, 0, (const uint8_t*)G->lit_new, gas->lit, G->n);
    if (hrr_total) {
        std::vector<double> part(nb);
        HIP_OK(hipMemcpyAsync(part.data(), G->cpart, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        double t = 0.0;
        for (double x : part) t += x;
        *hrr_total = t;
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_grid_set_cracks(phg_ctx* ctx, int grid, const uint8_t* lx, const uint8_t* ly, const uint8_t* lz) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return PHG_EINVAL;
        GridDev* G = get_grid(ctx, grid);
        if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
        HIP_OK(hipStreamSynchronize(ctx->stream));
        free_all(G->crack_allocs);
        G->crack = CrackDev();
        G->crack_nb = 0;
        G->crack_pre = nullptr;
        G->crack_q = nullptr;
        G->lx = G->ly = G->lz = nullptr;
        if (!lx && !ly && !lz) return PHG_OK;
        const int nx = G->nx, ny = G->ny, nz = G->nz, n = G->n;
        const size_t nf[3] = {(size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1)};
        const uint8_t* lk[3] = {lx, ly, lz};
        const float* dfo[3] = {G->fx, G->fy, G->fz};
        std::vector<uint8_t> blocked(n);
        HIP_OK(hipMemcpy(blocked.data(), G->blocked, n, hipMemcpyDeviceToHost));
        const float A = G->dx * G->dx;
        std::vector<int4> face;
        std::vector<int2> fzone;
        std::vector<float> oa;
        for (int ax = 0; ax < 3; ++ax) {
            if (!lk[ax]) continue;
            std::vector<float> fo(nf[ax]);
            HIP_OK(hipMemcpy(fo.data(), dfo[ax], nf[ax] * sizeof(float), hipMemcpyDeviceToHost));
            const int ex = nx + (ax == 0), ey = ny + (ax == 1);
            const int lim = ax == 0 ? nx : ax == 1 ? ny : nz, step = ax == 0 ? 1 : ax == 1 ? nx : nx * ny;
            for (size_t f = 0; f < nf[ax]; ++f) {
                if (!lk[ax][f]) continue;
                const int i = (int)(f % ex), j = (int)((f / ex) % ey), k = (int)(f / ((size_t)ex * ey));
                const int pos = ax == 0 ? i : ax == 1 ? j : k;
                const int cb = pos < lim ? i + nx * (j + ny * k) : -1;
                const int ca = pos > 0 ? i + nx * (j + ny * k) - step : -1;
                if ((ca >= 0 && blocked[ca]) || (cb >= 0 && blocked[cb])) continue;   // gasphase.cpp:424
                face.push_back(make_int4(ax, (int)f, ca, cb));
                fzone.push_back(make_int2(ca >= 0 ? G->zone[ca] : -1, cb >= 0 ? G->zone[cb] : -1));
                oa.push_back(fo[f] * A);
            }
        }
        const int ncr = (int)face.size();
        // per closed zone: the cracks of its root search (to another zone or
        // outside) and every crack that adds to its outflow, in crack order
        std::vector<int> kid(G->zone_open.size(), -1);
        int nclosed = 0;
        for (size_t z = 0; z < G->zone_open.size(); ++z)
            if (!G->zone_open[z]) kid[z] = nclosed++;
        std::vector<std::vector<int>> bl(nclosed), ol(nclosed);
        for (int q = 0; q < ncr; ++q) {
            const int za = fzone[q].x, zb = fzone[q].y;
            const int ka = za >= 0 ? kid[za] : -1, kb = zb >= 0 ? kid[zb] : -1;
            if (za != zb) {
                if (ka >= 0) bl[ka].push_back(q << 1 | 1);
                if (kb >= 0) bl[kb].push_back(q << 1);
            }
            if (ka >= 0) ol[ka].push_back(q << 1);
            if (kb >= 0) ol[kb].push_back(q << 1 | 1);
        }
        std::vector<int> bptr(nclosed + 1, 0), optr(nclosed + 1, 0), bent, oent;
        for (int k = 0; k < nclosed; ++k) {
            bent.insert(bent.end(), bl[k].begin(), bl[k].end());
            oent.insert(oent.end(), ol[k].begin(), ol[k].end());
            bptr[k + 1] = (int)bent.size();
            optr[k + 1] = (int)oent.size();
        }
        int rc = 0;
        auto up = [&](auto** dst, const auto* src, size_t cnt) {
            if (!rc) rc = upload(ctx, G->crack_allocs, dst, src, cnt);
        };
        uint8_t* dl[3] = {nullptr, nullptr, nullptr};
        for (int ax = 0; ax < 3; ++ax)
            if (lk[ax]) up(&dl[ax], lk[ax], nf[ax]);
        int4* dface = nullptr;
        int2* dfz = nullptr;
        float* doa = nullptr;
        int *dbp = nullptr, *dbe = nullptr, *dop = nullptr, *doe = nullptr;
        if (face.empty()) {
            face.push_back(make_int4(0, 0, -1, -1));
            fzone.push_back(make_int2(-1, -1));
            oa.push_back(0.0f);
        }
        if (bent.empty()) bent.push_back(0);
        if (oent.empty()) oent.push_back(0);
        up(&dface, face.data(), face.size());
        up(&dfz, fzone.data(), fzone.size());
        up(&doa, oa.data(), oa.size());
        up(&dbp, bptr.data(), bptr.size());
        up(&dbe, bent.data(), bent.size());
        up(&dop, optr.data(), optr.size());
        up(&doe, oent.data(), oent.size());
        if (!rc) rc = dalloc(ctx, G->crack_allocs, &G->crack_pre, (size_t)bptr[nclosed] + 1);
        if (!rc) rc = dalloc(ctx, G->crack_allocs, &G->crack_q, (size_t)ncr + 1);
        if (rc) {
            free_all(G->crack_allocs);
            return rc;
        }
        G->lx = dl[0];
        G->ly = dl[1];
        G->lz = dl[2];
        CrackDev& cd = G->crack;
        cd.ncr = ncr;
        cd.face = dface;
        cd.fzone = dfz;
        cd.oa = doa;
        cd.bptr = dbp;
        cd.bent = dbe;
        cd.optr = dop;
        cd.oent = doe;
        G->crack_nb = bptr[nclosed];
        return PHG_OK;
    });
}

int phg_grid_update_faces(phg_ctx* ctx, int grid, int n, const int* axis, const int* face, const float* f,
                          const float* a, const uint8_t* crack, float* u, float* v, float* w, int* rebuilt) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || n < 0 || (n > 0 && (!axis || !face || !f || !a || !crack))) return PHG_EINVAL;
        GridDev* G = get_grid(ctx, grid);
        if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
        if (rebuilt) *rebuilt = 0;
        const int nx = G->nx, ny = G->ny, nz = G->nz;
        const size_t nf[3] = {(size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1)};
        for (int q = 0; q < n; ++q) {
            if (axis[q] < 0 || axis[q] > 2 || face[q] < 0 || (size_t)face[q] >= nf[axis[q]])
                return fail(ctx, PHG_EINVAL, "update %d: no face %d on axis %d", q, face[q], axis[q]);
            if (!(f[q] >= 0 && f[q] <= 1) || !(a[q] >= 0 && std::isfinite(a[q])))
                return fail(ctx, PHG_EINVAL, "update %d: bad open fraction %g or coefficient %g", q, (double)f[q],
                            (double)a[q]);
        }
        if (n == 0) return PHG_OK;
        HIP_OK(hipSetDevice(ctx->device));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        // the geometry as the caller gave it: open fractions and solid mask
        // from the device, level 0's coefficients with the caller's values
        // restored on wind-closed sides, the crack flags
        std::vector<float> fo[3], co[3];
        std::vector<uint8_t> lk[3];
        const float* dfo[3] = {G->fx, G->fy, G->fz};
        const float* dco[3] = {G->ax, G->ay, G->az};
        const uint8_t* dlk[3] = {G->lx, G->ly, G->lz};
        bool had_cracks = false;
        for (int d = 0; d < 3; ++d) {
            fo[d].resize(nf[d]);
            co[d].resize(nf[d]);
            lk[d].assign(nf[d], 0);
            HIP_OK(hipMemcpy(fo[d].data(), dfo[d], nf[d] * sizeof(float), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(co[d].data(), dco[d], nf[d] * sizeof(float), hipMemcpyDeviceToHost));
            if (dlk[d]) {
                HIP_OK(hipMemcpy(lk[d].data(), dlk[d], nf[d], hipMemcpyDeviceToHost));
                had_cracks = true;
            }
        }
        for (int q = 0; q < 4; ++q) {
            if (!G->wc[q]) continue;
            size_t e = 0;
            const int side = q % 2;
            if (q < 2)
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j) co[0][(side ? nx : 0) + (nx + 1) * (j + ny * k)] = G->side_a[q][e++];
            else
                for (int k = 0; k < nz; ++k)
                    for (int i = 0; i < nx; ++i) co[1][i + nx * ((side ? ny : 0) + (ny + 1) * k)] = G->side_a[q][e++];
        }
        // apply; the matrix or the crack set changing is what marks the
        // reference's zones dirty (solver.cpp:98)
        bool matrix = false, cracks = false;
        std::vector<int2> closing;
        for (int q = 0; q < n; ++q) {
            const int d = axis[q];
            const size_t e = (size_t)face[q];
            const uint8_t c = crack[q] ? 1 : 0;
            matrix |= co[d][e] != a[q];
            cracks |= lk[d][e] != c || (c && fo[d][e] != f[q]);
            fo[d][e] = f[q];
            co[d][e] = a[q];
            lk[d][e] = c;
            if (f[q] == 0.0f) closing.push_back(make_int2(d, face[q]));
        }
        if (!closing.empty() && u && v && w) {
            std::vector<void*> tmp;   // freed on every path out of this block
            int2* dl = nullptr;
            int rz = upload(ctx, tmp, &dl, closing.data(), closing.size());
            if (!rz) {
                hipLaunchKernelGGL(k_face_zero, dim3(((int)closing.size() + TB - 1) / TB), dim3(TB), 0, ctx->stream,
                                   (const int2*)dl, (int)closing.size(), u, v, w);
                const hipError_t e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) rz = fail(ctx, PHG_EHIP, "zeroing closed faces: %s", hipGetErrorString(e));
            }
            free_all(tmp);
            if (rz) return rz;
        }
        bool any_crack = false;
        for (int d = 0; d < 3; ++d)
            for (uint8_t c : lk[d]) any_crack |= c != 0;
        auto set_cracks = [&](int id) -> int {
            if (!any_crack && !had_cracks && !cracks) return PHG_OK;
            if (!any_crack) return phg_grid_set_cracks(ctx, id, nullptr, nullptr, nullptr);
            return phg_grid_set_cracks(ctx, id, lk[0].data(), lk[1].data(), lk[2].data());
        };
        if (!matrix) {
            // zones and hierarchy stand: the open fractions and cracks only
            for (int d = 0; d < 3; ++d)
                HIP_OK(hipMemcpy((void*)dfo[d], fo[d].data(), nf[d] * sizeof(float), hipMemcpyHostToDevice));
            return cracks ? set_cracks(grid) : PHG_OK;
        }
        // the matrix changed: zones and the hierarchy are set up again
        // (computeZones, buildPressureLevels: gasphase.cpp:111-158,
        // pressure_mg.cpp:28-191) under the same id, precision, wind sides
        // and coarsest-solve choice
        std::vector<uint8_t> blocked(G->n);
        HIP_OK(hipMemcpy(blocked.data(), G->blocked, G->n, hipMemcpyDeviceToHost));
        // levels 0 and 1 (the work that grows with the cells) are formed on the
        // device from the wind-masked coefficients; the coarser levels, the
        // coarsest solve and the zones follow on the host.  If the device build
        // cannot take the grid, the whole set-up runs on the host.
        phg_grid_plan plan;
        plan.nx = nx;
        plan.ny = ny;
        plan.nz = nz;
        bool on_device = false;
        {
            std::vector<float> m[3] = {co[0], co[1], co[2]};
            for (int k = 0; k < nz; ++k) {
                for (int j = 0; j < ny; ++j) {
                    if (G->wc[0]) m[0][(nx + 1) * (j + ny * k)] = 0.0f;
                    if (G->wc[1]) m[0][nx + (nx + 1) * (j + ny * k)] = 0.0f;
                }
                for (int i = 0; i < nx; ++i) {
                    if (G->wc[2]) m[1][i + nx * ((ny + 1) * k)] = 0.0f;
                    if (G->wc[3]) m[1][i + nx * (ny + (ny + 1) * k)] = 0.0f;
                }
            }
            std::vector<void*> tmp;
            float* dm[3] = {nullptr, nullptr, nullptr};
            int ru = 0;
            for (int d = 0; d < 3 && !ru; ++d) ru = upload(ctx, tmp, &dm[d], m[d].data(), m[d].size());
            std::string derr = ru ? std::string("upload failed")
                                  : device_levels01(ctx->stream, nx, ny, nz, dm[0], dm[1], dm[2], G->blocked,
                                                    m[0].data(), m[1].data(), m[2].data(), plan.H);
            free_all(tmp);
            std::string herr;
            if (derr.empty() && finish_hierarchy(plan.H, kDenseMax, &herr)) {
                compute_zones(nx, ny, nz, co[0].data(), co[1].data(), co[2].data(), blocked.data(), plan.zone,
                              plan.zone_open);
                on_device = true;
            } else if (derr.empty()) {
                return fail(ctx, PHG_EINVAL, "%s", herr.c_str());
            }
        }
        GridDev* NG = nullptr;
        int rc = build_grid_dev(ctx, nx, ny, nz, G->dx, fo[0].data(), fo[1].data(), fo[2].data(), co[0].data(),
                                co[1].data(), co[2].data(), blocked.data(), G->has_wc ? G->wc : nullptr, G->prec,
                                on_device ? &plan : nullptr, &NG);
        if (rc) return rc;
        if (G->single_batch >= 0) {
            rc = phg_batch_destroy(ctx, G->single_batch);
            if (rc) {
                destroy_grid(NG);
                return rc;
            }
        }
        const int coarse = G->coarse, gen = G->gen;
        free_all(G->crack_allocs);
        free_all(G->allocs);
        *G = std::move(*NG);
        delete NG;
        G->gen = gen + 1;
        G->single_batch = -1;
        if (coarse == PHG_COARSE_GS20 && G->H.L.size() >= 2) G->coarse = PHG_COARSE_GS20;
        if (rebuilt) *rebuilt = on_device ? 2 : 1;
        return set_cracks(grid);
    });
}

// the leakage launches on the zone pressures in G->zp; the closed zones' rise
// rates are left in G->zdp
static int leakage_launch(phg_ctx* ctx, GridDev* G, float dt, const float* T, const float* Pb, const float* p, float* u,
                          float* v, float* w, float* S) {
    const size_t nzn = G->zone_open.size();
    const CrackDev& cd = G->crack;
    const int nc = G->nclosed;
    double* dint = G->zsum + G->nztask;
    if (nc) {
        const float V = G->dx * G->dx * G->dx;
        launch(ctx, K_LEAK, k_zone_sv, G->nztask, 0, (const int4*)G->ztask, (const int*)G->zcells, S, V, G->zsum);
        launch(ctx, K_LEAK, k_zone_fin_sum, 1, 0, nc, (const int*)G->ztptr, (const double*)G->zsum, dint);
    }
    if (cd.ncr) {
        HIP_OK(hipMemcpyAsync(G->zpn, G->zp, nzn * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        if (nc) {
            if (G->crack_nb)
                launch(ctx, K_LEAK, k_crack_prep, tiles(G->crack_nb), 0, cd, G->crack_nb, (const float*)G->zp, Pb, T,
                       G->crack_pre);
            launch(ctx, K_LEAK, k_crack_bisect, nc, 0, cd, (const int*)G->czone, (const float*)G->zvol,
                   (const double*)dint, (const float*)G->zp, (const float4*)G->crack_pre, dt, G->zpn);
        }
        launch(ctx, K_LEAK, k_crack_velocity, tiles(cd.ncr), 0, cd, (const float*)G->zpn, Pb, T, p, u, v, w,
               G->crack_q);
    }
    if (nc) {
        launch(ctx, K_LEAK, k_crack_zone_out, nc, 0, cd, (const int*)G->czone, (const float*)G->zvol,
               (const double*)dint, (const float*)G->zp, (const float*)G->crack_q, G->zdp, G->zval);
        launch(ctx, K_LEAK, k_zone_apply, tiles(G->n), 0, G->n, (const int*)G->zclosed, (const float*)G->zval, S, 1);
    }
    (void)nzn;
    return PHG_OK;
}

int phg_zone_leakage(phg_ctx* ctx, int grid, float dt, const float* zoneP, const float* T, const float* Pb,
                     const float* p, float* u, float* v, float* w, float* S, float* zoneDP, int* leak_faces) {
    if (!ctx || !zoneP || !T || !Pb || !p || !u || !v || !w || !S) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (!(dt > 0)) return fail(ctx, PHG_EINVAL, "dt must be positive");
    const size_t nzn = G->zone_open.size();
    const CrackDev& cd = G->crack;
    if (leak_faces) *leak_faces = cd.ncr;
    if (zoneDP) std::fill(zoneDP, zoneDP + nzn, 0.0f);
    if (nzn == 0 || (G->nclosed == 0 && cd.ncr == 0)) return PHG_OK;   // nothing to shift, no crack flow
    HIP_OK(hipMemcpyAsync(G->zp, zoneP, nzn * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    const int nc = G->nclosed;
    if (int rc = leakage_launch(ctx, G, dt, T, Pb, p, u, v, w, S)) return rc;
    if (nc) {
        std::vector<float> zd(nc);
        HIP_OK(hipMemcpyAsync(zd.data(), G->zdp, nc * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        if (zoneDP)
            for (int k = 0, z = 0; z < (int)nzn; ++z)
                if (!G->zone_open[z]) zoneDP[z] = zd[k++];
    } else {
        // zoneP's staging copy must land before the caller may reuse it
        HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_zone_pressures_set(phg_ctx* ctx, int grid, const float* zoneP) {
    if (!ctx || !zoneP) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const size_t nzn = G->zone_open.size();
    if (nzn == 0) return PHG_OK;
    HIP_OK(hipMemcpyAsync(G->zp, zoneP, nzn * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemsetAsync(G->zdp, 0, ((size_t)G->nclosed + 1) * sizeof(float), ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));   // the caller's array may go at once
    return PHG_OK;
}

int phg_zone_pressures_get(phg_ctx* ctx, int grid, float* zoneP, float* zoneDP) {
    if (!ctx || !zoneP) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const size_t nzn = G->zone_open.size();
    if (nzn == 0) return PHG_OK;
    std::vector<float> zd((size_t)G->nclosed + 1);
    HIP_OK(hipMemcpyAsync(zoneP, G->zp, nzn * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(zd.data(), G->zdp, zd.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (zoneDP)
        for (size_t k = 0, z = 0; z < nzn; ++z) zoneDP[z] = G->zone_open[z] ? 0.0f : zd[k++];
    return PHG_OK;
}

int phg_zone_leakage_resident(phg_ctx* ctx, int grid, float dt, const float* T, const float* Pb, const float* p,
                              float* u, float* v, float* w, float* S) {
    if (!ctx || !T || !Pb || !p || !u || !v || !w || !S) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (!(dt > 0)) return fail(ctx, PHG_EINVAL, "dt must be positive");
    if (G->zone_open.empty() || (G->nclosed == 0 && G->crack.ncr == 0)) return PHG_OK;
    if (int rc = leakage_launch(ctx, G, dt, T, Pb, p, u, v, w, S)) return rc;
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_zone_pressure_resident(phg_ctx* ctx, int grid, float dt, float* Pb) {
    if (!ctx || !Pb) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const int nc = G->nclosed;
    if (nc == 0) return PHG_OK;
    launch(ctx, K_LEAK, k_zone_pressure, (nc + TB - 1) / TB, 0, nc, (const int*)G->czone, (const float*)G->zdp, dt,
           G->zp, G->zval);
    launch(ctx, K_ZONE, k_zone_apply, tiles(G->n), 0, G->n, (const int*)G->zclosed, (const float*)G->zval, Pb, 0);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_grid_stats(phg_ctx* ctx, int grid, float* max_vel, float* max_nu, float* max_div) {
    if (!ctx) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    float v[3];
    HIP_OK(hipMemcpyAsync(v, G->vmax, sizeof v, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (max_vel) *max_vel = v[0];
    if (max_nu) *max_nu = v[1];
    if (max_div) *max_div = v[2];
    return PHG_OK;
}

int phg_zone_integrals(phg_ctx* ctx, int grid, const float* S, double* dint) {
    if (!ctx || !S || !dint) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    const size_t nzn = G->zone_open.size();
    std::fill(dint, dint + nzn, 0.0);
    if (G->nclosed == 0) return PHG_OK;
    const float V = G->dx * G->dx * G->dx;
    launch(ctx, K_ZONE, k_zone_sv, G->nztask, 0, (const int4*)G->ztask, (const int*)G->zcells, S, V, G->zsum);
    launch(ctx, K_ZONE, k_zone_fin_sum, 1, 0, G->nclosed, (const int*)G->ztptr, (const double*)G->zsum,
           G->zsum + G->nztask);
    std::vector<double> zs(G->nclosed);
    HIP_OK(hipMemcpyAsync(zs.data(), G->zsum + G->nztask, G->nclosed * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    int k = 0;
    for (size_t z = 0; z < nzn; ++z)
        if (!G->zone_open[z]) dint[z] = zs[k++];
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_zone_values(phg_ctx* ctx, int grid, const float* values, float* x, int subtract) {
    if (!ctx || !values || !x) return PHG_EINVAL;
    GridDev* G = get_grid(ctx, grid);
    if (!G) return fail(ctx, PHG_ESTATE, "unknown grid %d", grid);
    if (G->nclosed == 0) return PHG_OK;
    std::vector<float> cv;
    cv.reserve(G->nclosed);
    for (size_t z = 0; z < G->zone_open.size(); ++z)
        if (!G->zone_open[z]) cv.push_back(values[z]);
    HIP_OK(hipMemcpyAsync(G->zval, cv.data(), cv.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    launch(ctx, K_ZONE, k_zone_apply, tiles(G->n), 0, G->n, (const int*)G->zclosed, (const float*)G->zval, x,
           subtract ? 1 : 0);
    // the staging vector is pageable: wait before it goes out of scope
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_copy(phg_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src || bytes % 16) return ctx ? fail(ctx, PHG_EINVAL, "bytes must be a multiple of 16") : PHG_EINVAL;
    const long n = (long)(bytes / 16);
    // one element per thread: a grid-stride loop over a few thousand blocks
    // moved 4.6-4.9 TB/s on MI355X, this form 6.2-6.3 TB/s
    if ((n + TB - 1) / TB > INT32_MAX) return fail(ctx, PHG_EINVAL, "copy too large");
    const int blocks = (int)std::max<long>(1, (n + TB - 1) / TB);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(TB), 0, ctx->stream, (const float4*)src, (float4*)dst, n);
    HIP_OK(hipGetLastError());
    return PHG_OK;
}

int phg_profile_enable(phg_ctx* ctx, int on) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    prof_drain(ctx);
    ctx->prof = on != 0;
    return PHG_OK;
}

int phg_profile_select(phg_ctx* ctx, unsigned long long class_mask) {
    if (!ctx) return PHG_EINVAL;
    ctx->prof_mask = class_mask;
    return PHG_OK;
}

int phg_profile_get(phg_ctx* ctx, int index, const char** name, double* total_ms, long long* launches, double* bytes) {
    if (!ctx || index < 0 || index >= K_NCLASS) return PHG_EINVAL;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    prof_drain(ctx);
    if (name) *name = kKName[index];
    if (total_ms) *total_ms = ctx->prof_ms[index];
    if (launches) *launches = ctx->prof_n[index];
    if (bytes) *bytes = 0;
    return PHG_OK;
}

int phg_profile_reset(phg_ctx* ctx) {
    if (!ctx) return PHG_EINVAL;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    prof_drain(ctx);
    for (int i = 0; i < K_NCLASS; ++i) {
        ctx->prof_ms[i] = 0;
        ctx->prof_n[i] = 0;
    }
    return PHG_OK;
}

}  // extern "C"
