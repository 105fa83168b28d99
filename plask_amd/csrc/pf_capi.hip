// C ABI of the gfx950 FEM solve path (declarations and reference citations in
// include/plask_fem.h).  Owns device buffers, builds the eliminated system
// matrix-free and drives the Jacobi-preconditioned CG: one fused
// apply + (p,Ap) kernel and one vector-update + (r,z) kernel per iteration,
// scalars kept on the device, chunks of iterations replayed from a hipGraph,
// and the host reading back a 64-byte state once per chunk.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/plask_fem.h"
#include "fem_kernels.hpp"
#include "batch_kernels.hpp"
#include "mg_kernels.hpp"

using namespace pf;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

// One coarse level of the semicoarsening multigrid preconditioner: the in-plane axes
// coarsened (3D: axes 0 and 1; 2D: axis 0, with n1 = 1), all planes of the layer axis
// kept.  Device arrays; the level above is the next finer one (level 0 is the solve
// grid itself).
struct MgLevel {
    int n0 = 0, n1 = 0;
    long long N = 0, E = 0;
    double *hx = nullptr, *hy = nullptr, *rmid = nullptr;  // hy: 3D only; rmid: 2D cylindrical
    float *hx32 = nullptr, *hy32 = nullptr, *rmid32 = nullptr;
    double *kh = nullptr, *kv = nullptr;
    float *kh32 = nullptr, *kv32 = nullptr;
    uint64_t* dmask = nullptr;
    // transfer from the finer level (tables described in mg_kernels.hpp)
    int *cx = nullptr, *cy = nullptr, *mx = nullptr, *my = nullptr;
    double *wx = nullptr, *wy = nullptr;
    float *wx32 = nullptr, *wy32 = nullptr;
    // restriction rows per coarse node of each axis: 3 fine indices and weights
    int *rx = nullptr, *ry = nullptr;
    double *rwx = nullptr, *rwy = nullptr;
    float *rwx32 = nullptr, *rwy32 = nullptr;
    // partitioned line factors (same partition lengths as the solve grid)
    double *la = nullptr, *lw = nullptr, *lcp = nullptr, *schur = nullptr;
    float *la32 = nullptr, *lw32 = nullptr, *lcp32 = nullptr, *schur32 = nullptr;
    int schur_m64 = -1, schur_m32 = -1;
    int pm64 = 0, pm32 = 0;  // partition lengths of this level's line solve
    // right-hand side, iterate and residual of the level
    double *b = nullptr, *x = nullptr, *r = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr;
    // transient operator: the level's capacity diagonal (P^T of the finer level's)
    double* dsh = nullptr;
    float* dsh32 = nullptr;
};

struct pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int dim = 0, geom = 0, equation = PF_THERMAL;
    int64_t n[3] = {1, 1, 1};
    int64_t N = 0, E = 0;
    std::vector<double> ax[3];

    // geometry (FP64 and FP32 copies)
    double *hx = nullptr, *hy = nullptr, *hz = nullptr, *rmid = nullptr;
    float *hx32 = nullptr, *hy32 = nullptr, *hz32 = nullptr, *rmid32 = nullptr;
    // element data
    double *kh = nullptr, *kv = nullptr, *q = nullptr;
    float *kh32 = nullptr, *kv32 = nullptr;
    bool have_coeffs = false, have_source = false;
    // Dirichlet
    uint64_t* dmask = nullptr;
    long long* dnodes = nullptr;
    double* dvals = nullptr;
    int64_t nd = 0;
    double dvals_sq = 0.;
    // line-factor set-up scratch: per-node line coefficients, per-segment sums
    double *sb = nullptr, *sc = nullptr, *sff = nullptr, *sfl = nullptr;
    size_t nseg_cap = 0;  // entries of sff / sfl
    // interpolation workspace (axes, points, results)
    double *iax = nullptr, *ipts = nullptr, *iout = nullptr, *imid = nullptr;
    size_t ipts_cap = 0, iout_cap = 0;
    // nodal load added to the element load vector (boundary heat flux / radiation)
    long long* lnodes = nullptr;
    double* lvals = nullptr;
    int64_t nl = 0;
    // vectors
    double *p2 = nullptr;
    float *p2_32 = nullptr;
    // line preconditioner factors (a, w, c')
    double *la = nullptr, *lw = nullptr, *lcp = nullptr;
    float *la32 = nullptr, *lw32 = nullptr, *lcp32 = nullptr;
    double *x = nullptr, *r = nullptr, *p = nullptr, *qv = nullptr, *z = nullptr, *dinv = nullptr, *F = nullptr,
           *tmp = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *z32 = nullptr, *dinv32 = nullptr;
    PcgState* st = nullptr;
    PcgState* hst = nullptr;  // pinned host mirror
    double* partials = nullptr;
    size_t npartials = 0;
    bool dirty = true;  // F, Dinv, ||B|| need recomputing
    double bnorm = 0.;
    bool have_solution = false;

    // launch configuration
    int BY[2] = {4, 4}, KC[2] = {8, 16}, KC2 = 16;  // [0] FP64, [1] FP32
    int pfd = 1;  // planes prefetched by the 3D apply (1 or 2)
    // 3D apply per precision ([0] FP64, [1] FP32): 0 tile blocks (BY x KC), 1 row blocks (KR planes)
    int apply_kind[2] = {0, 0}, KR[2] = {16, 16};
    bool graphs = true;  // replay CG chunks from hipGraphs
    int xcd_order = 1;   // tile apply: XCD-aware block order
    // line solve: partition length per precision (0: sequential whole-line kernel)
    int line_m = 0;  // requested (0: default 8 FP64 / 16 FP32, < 0: sequential)
    int pm64 = 8, pm32 = 16;
    double* schur = nullptr;  // Thomas factors of the separator systems
    float* schur32 = nullptr;
    int upd_blocks = 1024;
    // multigrid preconditioner (3D): coarse levels, smoother residual of level 0, tuning
    std::vector<MgLevel> mg;
    bool mg_built = false;
    double *mg_r0 = nullptr;
    float *mg_r0_32 = nullptr;
    int mg_max_levels = 0;      // 0: coarsen until both in-plane axes have <= 3 nodes
    int mg_coarse_sweeps = 1;   // smoothing sweeps on the coarsest level
    int mg_xcd = 1;             // multigrid levels >= 1: XCD-aligned residual tiles and sweep groups
    int mg_fuse = 0;            // 3D: grid transfers fused into the smoother / residual kernels (measured slower)
    double mg_omega = 0.65;     // damping of the line-Jacobi smoother
    // cached graphs of one chunk of iterations
    hipGraphExec_t graph64 = nullptr, graph32 = nullptr;
    int graph_chunk[2] = {0, 0}, graph_key[2] = {-1, -1};  // what graph64 / graph32 hold
    bool prec32 = false;  // FP64 PCG with the preconditioner applied in FP32 (PF_FP64_PREC32)

    // profiling: HIP event pairs around kernels of profiled CG chunks, summed per
    // category (PF_KCAT_*) over the iterations that ran
    bool profile = false;
    int prof_iter = -1;  // iteration index inside a profiled chunk, -1 outside
    // cat: a kernel class (PF_KCAT_*, -1 none); kern / level / tbytes / bytes: one launch
    // of kernel PF_K_* on multigrid level `level` in precision tbytes (8 or 4) moving
    // `bytes` algorithmic HBM bytes (kern -1: a class record only)
    struct ProfRec {
        int a, b, cat, iter, kern, level, tbytes;
        double bytes;
        bool single;  // the events time one kernel's execution (set by its launch)
    };
    std::vector<hipEvent_t> prof_ev;
    std::vector<ProfRec> prof_recs;
    size_t prof_used = 0;
    int prof_pending = -1;  // kernel record whose events the next pf_launch takes
    double kt[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    int64_t kc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<pf_kernel_stat> kstat;  // per (kernel, level, precision) of the profiled solve

    // transient (theta-method) operator A = theta K + C: the lumped capacity diagonal C
    // (FP64 / FP32), the stiffness K as set (khB, kvB; kh, kv then hold theta K), the
    // static load F0, the B T^n terms of the Dirichlet nodes, an all-free mask
    bool dyn = false;
    double theta = 1.;
    size_t ydir_cap = 0;
    double *dcap = nullptr, *khB = nullptr, *kvB = nullptr, *F0 = nullptr, *ydir = nullptr;
    uint64_t* zmask = nullptr;
    // diagonal shift of the operator (dsh = dcap + dconv, FP64 / FP32), on while either is set
    bool dsh_on = false;
    double* dsh = nullptr;
    float* dsh32 = nullptr;
    // boundary matrix (pf_boundary_matrix_set): its diagonal dconv, its off-diagonal part in
    // CSR over the nbr rows it touches
    bool have_dconv = false, have_bnd = false;
    double* dconv = nullptr;
    long long nbr = 0;
    long long *bnd_rows = nullptr, *bnd_cols = nullptr;
    int* bnd_ptr = nullptr;
    double* bnd_vals = nullptr;
    float* bnd_vals32 = nullptr;

    // Picard-loop state
    double *k0h = nullptr, *k0v = nullptr, *ah = nullptr, *av = nullptr;
    bool have_law = false;
    double* xprev = nullptr;
    double *cur = nullptr, *heat = nullptr, *emean = nullptr;
    double *flux = nullptr, *kpost = nullptr;  // element heat flux, conductivities given for it
    bool have_cur = false, have_flux = false;
    unsigned char *sel = nullptr, *noheat = nullptr;
    bool have_sel = false, have_noheat = false;
    double* dout = nullptr;  // small device results
    long long *je = nullptr, *jbot = nullptr, *jtop = nullptr;
    int* jid = nullptr;
    double *jheight = nullptr, *jbeta = nullptr, *jjs = nullptr;
    int64_t nje = 0;
    int jcorner = 0;

    std::vector<void*> allocs;
};

#define HIPCK(call)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                            \
            return PF_EDEVICE;                                                                       \
        }                                                                                            \
    } while (0)

static int fail(pf_ctx* ctx, int code, const std::string& msg) {
    ctx->err = msg;
    return code;
}

template <typename T> static int dalloc(pf_ctx* ctx, T** ptr, size_t count) {
    if (*ptr) {
        hipFree(*ptr);
        auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), (void*)*ptr);
        if (it != ctx->allocs.end()) ctx->allocs.erase(it);
        *ptr = nullptr;
    }
    if (count == 0) return PF_OK;
    HIPCK(hipMalloc((void**)ptr, sizeof(T) * count));
    ctx->allocs.push_back((void*)*ptr);
    return PF_OK;
}

static void drop_graphs(pf_ctx* ctx) {
    if (ctx->graph64) hipGraphExecDestroy(ctx->graph64);
    if (ctx->graph32) hipGraphExecDestroy(ctx->graph32);
    ctx->graph64 = ctx->graph32 = nullptr;
    ctx->graph_key[0] = ctx->graph_key[1] = -1;
}

// Profiled chunks: prof_open records a start event for one kernel category (returns a
// handle, -1 when not profiling), prof_close its end event; prof_collect (after a stream
// synchronisation) adds the durations of the records of iterations < ran.
static int prof_rec(pf_ctx* ctx, int cat, int kern, int level, int tbytes, double bytes, bool single = false) {
    if (!ctx->profile || ctx->prof_iter < 0) return -1;
    while (ctx->prof_used + 2 > ctx->prof_ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->prof_ev.push_back(e);
    }
    pf_ctx::ProfRec r{(int)ctx->prof_used, (int)ctx->prof_used + 1, cat, ctx->prof_iter, kern, level, tbytes, bytes,
                      single};
    ctx->prof_used += 2;
    if (!single) hipEventRecord(ctx->prof_ev[r.a], ctx->stream);
    ctx->prof_recs.push_back(r);
    return (int)ctx->prof_recs.size() - 1;
}

static int prof_open(pf_ctx* ctx, int cat) { return prof_rec(ctx, cat, -1, 0, 0, 0.); }

// one launch of kernel `kern` (PF_K_*) on level `level` in precision T moving `bytes`.
// single: the kernel is the next pf_launch, whose events then hold its start and end of
// execution (what a kernel trace reports); else events bracket the launches on the stream
// until prof_close (which adds the dispatch overhead of each launch).
template <typename T> static int prof_kern(pf_ctx* ctx, int kern, int level, double bytes, bool single = true) {
    const int h = prof_rec(ctx, -1, kern, level, (int)sizeof(T), bytes, single);
    if (h >= 0 && single) ctx->prof_pending = h;
    return h;
}

// Kernel launch on `stream`; takes the events of a pending kernel record (prof_kern).
template <typename K, typename... A>
static inline void pf_launch(pf_ctx* ctx, K kernel, dim3 grid, dim3 block, size_t shm, hipStream_t stream, A... args) {
    const int h = ctx->prof_pending;
    if (h >= 0) {
        ctx->prof_pending = -1;
        const pf_ctx::ProfRec& r = ctx->prof_recs[h];
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shm, stream, ctx->prof_ev[r.a], ctx->prof_ev[r.b], 0u,
                              args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, stream, args...);
    }
}

// Algorithmic HBM bytes of the launches (each value read or written once; element data
// two conductivities; the Dirichlet mask one bit per node, whole 64-bit words).
static double mask_bytes(long long n) { return 8.0 * (double)((n + 63) / 64); }
// fused PCG apply: z, p_old, x read; p_new, q, x written
static double bytes_apply_fused(long long n, long long e, int t) { return t * (6.0 * n + 2.0 * e) + mask_bytes(n); }
// residual r = b - A x: x, b read; r written
static double bytes_residual(long long n, long long e, int t) { return t * (3.0 * n + 2.0 * e) + mask_bytes(n); }
// damped line sweep: right-hand side and the three line factors read, the sweep written;
// with the PCG update also q read and the right-hand side written back; a previous
// iterate read when there is one; the PCG residual read for the final dot products;
// three factors per separator and column
static double bytes_smooth(long long n, int ncol, int P, int t, bool upd, bool xin, bool fin) {
    const double v = 5.0 + (upd ? 2.0 : 0.0) + (xin ? 1.0 : 0.0) + (fin && !upd ? 1.0 : 0.0);
    return t * (v * n + 3.0 * std::max(P - 1, 0) * (double)ncol);
}

static void prof_close(pf_ctx* ctx, int h) {
    if (h < 0) return;
    const pf_ctx::ProfRec& r = ctx->prof_recs[h];
    if (ctx->prof_pending == h) {  // no launch took the record: an empty interval
        ctx->prof_pending = -1;
        hipEventRecord(ctx->prof_ev[r.a], ctx->stream);
        hipEventRecord(ctx->prof_ev[r.b], ctx->stream);
    } else if (!r.single) {
        hipEventRecord(ctx->prof_ev[r.b], ctx->stream);
    }
}

static void prof_collect(pf_ctx* ctx, int ran) {
    for (const auto& r : ctx->prof_recs) {
        if (r.iter >= ran) continue;
        float t = 0.f;
        hipEventElapsedTime(&t, ctx->prof_ev[r.a], ctx->prof_ev[r.b]);
        if (r.cat >= 0) {
            ctx->kt[r.cat] += 1e-3 * t;
            ctx->kc[r.cat] += 1;
        }
        if (r.kern >= 0) {
            pf_kernel_stat* k = nullptr;
            for (auto& q : ctx->kstat)
                if (q.kernel == r.kern && q.level == r.level && q.value_bytes == r.tbytes) k = &q;
            if (!k) {
                ctx->kstat.push_back(pf_kernel_stat{r.kern, r.level, r.tbytes, 0, 0., 0.});
                k = &ctx->kstat.back();
            }
            k->launches += 1;
            k->seconds += 1e-3 * t;
            k->bytes += r.bytes;
        }
    }
    ctx->prof_recs.clear();
    ctx->prof_used = 0;
}

static void dfree(pf_ctx* ctx, void** ptr) {
    if (!*ptr) return;
    hipFree(*ptr);
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), *ptr);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    *ptr = nullptr;
}

static void dyn_capture(pf_ctx* ctx);

// release the multigrid levels (grid change)
static void mg_free(pf_ctx* ctx) {
    for (MgLevel& L : ctx->mg) {
        void** ptrs[] = {(void**)&L.hx,    (void**)&L.hy,    (void**)&L.hx32,    (void**)&L.hy32,  (void**)&L.kh,
                         (void**)&L.kv,    (void**)&L.kh32,  (void**)&L.kv32,    (void**)&L.dmask, (void**)&L.cx,
                         (void**)&L.cy,    (void**)&L.mx,    (void**)&L.my,      (void**)&L.wx,    (void**)&L.wy,
                         (void**)&L.wx32,  (void**)&L.wy32,  (void**)&L.la,      (void**)&L.lw,    (void**)&L.lcp,
                         (void**)&L.schur, (void**)&L.la32,  (void**)&L.lw32,    (void**)&L.lcp32, (void**)&L.schur32,
                         (void**)&L.b,     (void**)&L.x,     (void**)&L.r,       (void**)&L.b32,   (void**)&L.x32,
                         (void**)&L.r32,   (void**)&L.rmid,  (void**)&L.rmid32, (void**)&L.dsh,
                         (void**)&L.dsh32, (void**)&L.rx,    (void**)&L.ry,     (void**)&L.rwx,
                         (void**)&L.rwy,   (void**)&L.rwx32, (void**)&L.rwy32};
        for (void** q : ptrs) dfree(ctx, q);
    }
    ctx->mg.clear();
    dfree(ctx, (void**)&ctx->mg_r0);
    dfree(ctx, (void**)&ctx->mg_r0_32);
    ctx->mg_built = false;
}

// Launch shape of the partitioned line kernel: P segments of m planes per column,
// CB columns per block (CB * P threads, at most PartLB).
struct PartShape {
    int m = 0, P = 1, CB = 64, threads = 64;
};

template <typename T>
static int part_lb(int m) {
    switch (m) {
        case 4: return PartLB<T, 4>::value;
        case 8: return PartLB<T, 8>::value;
        case 16: return PartLB<T, 16>::value;
        case 32: return PartLB<T, 32>::value;
    }
    return 0;
}

template <typename T>
static PartShape part_shape(int m, int n, int ncol) {
    PartShape ps;
    if (m <= 0) return ps;
    const int lb = part_lb<T>(m);
    ps.m = m;
    ps.P = (n + m - 1) / m;
    if (ps.P > lb) {  // line too long for one block: caller falls back
        ps.m = 0;
        return ps;
    }
    ps.CB = 64;
    while (ps.CB > 1 && ps.CB * ps.P > lb) ps.CB >>= 1;
    while (ps.CB > 1 && ps.CB / 2 >= ncol) ps.CB >>= 1;
    ps.threads = (ps.CB * ps.P + 63) / 64 * 64;
    return ps;
}

// Partition lengths of the line solve and the separator-factor buffers.  A line
// longer than one block can hold (P > threads per block) uses the sequential
// whole-line kernel for that precision.
static int line_config(pf_ctx* ctx) {
    const int ncol = (int)(ctx->dim == 3 ? ctx->n[0] * ctx->n[1] : ctx->n[0]);
    const int n = (int)(ctx->dim == 3 ? ctx->n[2] : ctx->n[1]);
    int m64 = 8, m32 = 16;
    if (ctx->line_m < 0) m64 = m32 = 0;
    else if (ctx->line_m > 0) {
        m32 = ctx->line_m;
        m64 = std::min(ctx->line_m, 16);
    }
    PartShape s64 = part_shape<double>(m64, n, ncol), s32 = part_shape<float>(m32, n, ncol);
    for (PartShape* ps : {&s64, &s32})
        if (ps->m && 2 * (size_t)((ncol + ps->CB - 1) / ps->CB) > ctx->npartials) ps->m = 0;
    ctx->pm64 = s64.m;
    ctx->pm32 = s32.m;
    int rc;
    const size_t n64 = s64.m ? 3 * (size_t)std::max(s64.P - 1, 1) * ncol : 0;
    const size_t n32 = s32.m ? 3 * (size_t)std::max(s32.P - 1, 1) * ncol : 0;
    if ((rc = dalloc(ctx, &ctx->schur, n64))) return rc;
    if ((rc = dalloc(ctx, &ctx->schur32, n32))) return rc;
    // set-up scratch (the coarse levels have the same planes and fewer columns)
    const int pmin = (s64.m && s32.m) ? std::min(s64.m, s32.m) : std::max(s64.m, s32.m);
    const size_t nseg = pmin ? (size_t)((n + pmin - 1) / pmin) * ncol : 0;
    if ((rc = dalloc(ctx, &ctx->sb, nseg ? (size_t)ctx->N : 0)) || (rc = dalloc(ctx, &ctx->sc, nseg ? (size_t)ctx->N : 0)) ||
        (rc = dalloc(ctx, &ctx->sff, nseg)) || (rc = dalloc(ctx, &ctx->sfl, nseg)))
        return rc;
    ctx->nseg_cap = nseg;
    ctx->dirty = true;  // the factors depend on the partition
    drop_graphs(ctx);
    return PF_OK;
}

static inline int grid1(long long n, int bs = 256, int maxb = 2048) {
    long long b = (n + bs - 1) / bs;
    return (int)std::max(1LL, std::min<long long>(b, maxb));
}

// Workspace of pf_solution_prolong_dev, the entry point that allocates nothing: the
// per-axis interpolation tables (an int index and a float weight per fine node of each
// axis) live in this file-scope device array.  Each GPU holds its own copy until the
// process ends; host code uses the address hipGetSymbolAddress returns for the device
// (scratch_ptr[device], set when a context is created on it), never the array's name.
#define PF_PROLONG_SCRATCH_BYTES (256 * 1024)
static __device__ __align__(256) unsigned char scratch[PF_PROLONG_SCRATCH_BYTES];
static void* scratch_ptr[64] = {nullptr};
static std::mutex scratch_mu;

static hipError_t scratch_init(int device) {
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // the current device's copy
    if (e == hipSuccess) scratch_ptr[device] = ptr;
    return e;
}

// A kernel-tracing profiler attached to the process (rocprofv3 exports its tool
// library list and output path into the profiled program's environment).  Under
// rocprofv3 --kernel-trace, replaying hipGraphs ended the process in a host-side copy
// inside the profiler library reading a device address (the graph's kernel arguments);
// the same solves with plain launches trace cleanly.  Contexts created under such a
// profiler therefore start with graphs off (pf_set_graphs can turn them back on).
static bool profiler_attached() {
    const char* vars[] = {"ROCP_TOOL_LIBRARIES", "ROCPROF_OUTPUT_PATH"};
    for (const char* v : vars) {
        const char* s = std::getenv(v);
        if (s && *s) return true;
    }
    return false;
}

static bool graphs_default(const char* who) {
    static bool warned = false;
    if (!profiler_attached()) return true;
    if (!warned) {
        std::fprintf(stderr, "%s: profiler attached (ROCP_TOOL_LIBRARIES / ROCPROF_OUTPUT_PATH set): "
                             "hipGraph replay off, kernels launched one by one\n", who);
        warned = true;
    }
    return false;
}

extern "C" {

int pf_ctx_create(pf_ctx** out, int device, int flags) {
    if (!out) return PF_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PF_EDEVICE;
    if (device < 0 || device >= ndev) return PF_EINVAL;
    pf_ctx* ctx = new pf_ctx();
    ctx->device = device;
    ctx->profile = (flags & 1) != 0;
    ctx->graphs = graphs_default("plask_amd");
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PF_EDEVICE;
    }
    if (hipMalloc((void**)&ctx->st, sizeof(PcgState)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->hst, sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return PF_EDEVICE;
    }
    hipMemset(ctx->st, 0, sizeof(PcgState));
    if (scratch_init(device) != hipSuccess) {  // device current (hipSetDevice above)
        pf_ctx_destroy(ctx);
        return PF_EDEVICE;
    }
    *out = ctx;
    return PF_OK;
}

void pf_ctx_destroy(pf_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    for (hipEvent_t e : ctx->prof_ev) hipEventDestroy(e);
    for (void* p : ctx->allocs) hipFree(p);
    if (ctx->st) hipFree(ctx->st);
    if (ctx->hst) hipHostFree(ctx->hst);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* pf_last_error(const pf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t pf_num_nodes(const pf_ctx* ctx) { return ctx ? ctx->N : 0; }
int64_t pf_num_elements(const pf_ctx* ctx) { return ctx ? ctx->E : 0; }
void* pf_solution_device(pf_ctx* ctx) { return ctx ? (void*)ctx->x : nullptr; }

// LDS bytes of the FP64 3D tile kernel: two x tiles of (rows + 1) x (n0 + 1) and two
// corner-output buffers of 4 x rows x (n0 + 1)
static size_t tile_lds(int rows, int n0) {
    return sizeof(double) * (size_t)(2 * (rows + 1) + 8 * rows) * (size_t)(n0 + 1);
}

int pf_grid_set(pf_ctx* ctx, int dim, int geom, const double* ax0, int64_t n0, const double* ax1, int64_t n1,
                const double* ax2, int64_t n2) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    if (dim != 2 && dim != 3) return fail(ctx, PF_EINVAL, "dim must be 2 or 3");
    if (geom == PF_GEOM_CYLINDRICAL && dim != 2) return fail(ctx, PF_EINVAL, "cylindrical geometry is 2D only");
    if (!ax0 || !ax1 || n0 < 2 || n1 < 2 || (dim == 3 && (!ax2 || n2 < 2)))
        return fail(ctx, PF_EINVAL, "each axis needs at least 2 nodes");
    if (n0 > 1024) return fail(ctx, PF_ESHAPE, "axis 0 longer than 1024 nodes is not supported by the row kernels");
    if (dim == 3 && n0 > 512) return fail(ctx, PF_ESHAPE, "3D: axis 0 longer than 512 nodes is not supported");
    const double* axs[3] = {ax0, ax1, ax2};
    int64_t ns[3] = {n0, n1, dim == 3 ? n2 : 1};
    for (int a = 0; a < dim; ++a) {
        for (int64_t i = 0; i + 1 < ns[a]; ++i)
            if (!(axs[a][i + 1] > axs[a][i])) return fail(ctx, PF_EINVAL, "axis coordinates must increase strictly");
    }
    if (geom == PF_GEOM_CYLINDRICAL && ax0[0] < 0.) return fail(ctx, PF_EINVAL, "radius must be non-negative");
    drop_graphs(ctx);
    mg_free(ctx);
    ctx->dim = dim;
    ctx->geom = geom;
    for (int a = 0; a < 3; ++a) {
        ctx->n[a] = ns[a];
        ctx->ax[a].assign(axs[a] ? axs[a] : nullptr, axs[a] ? axs[a] + ns[a] : nullptr);
    }
    ctx->N = ns[0] * ns[1] * ns[2];
    ctx->E = (ns[0] - 1) * (ns[1] - 1) * (dim == 3 ? ns[2] - 1 : 1);
    // element sizes
    double** hd[3] = {&ctx->hx, &ctx->hy, &ctx->hz};
    float** hf[3] = {&ctx->hx32, &ctx->hy32, &ctx->hz32};
    for (int a = 0; a < dim; ++a) {
        std::vector<double> h(ns[a] - 1);
        std::vector<float> hf32(ns[a] - 1);
        for (int64_t i = 0; i + 1 < ns[a]; ++i) {
            h[i] = axs[a][i + 1] - axs[a][i];
            hf32[i] = (float)h[i];
        }
        int rc;
        if ((rc = dalloc(ctx, hd[a], h.size()))) return rc;
        if ((rc = dalloc(ctx, hf[a], h.size()))) return rc;
        HIPCK(hipMemcpy(*hd[a], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(*hf[a], hf32.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    }
    if (geom == PF_GEOM_CYLINDRICAL) {
        std::vector<double> rm(n0 - 1);
        std::vector<float> rm32(n0 - 1);
        for (int64_t i = 0; i + 1 < n0; ++i) {
            rm[i] = 0.5 * (ax0[i] + ax0[i + 1]);  // midpoint.rad_r() (therm2d.cpp:333)
            rm32[i] = (float)rm[i];
        }
        int rc;
        if ((rc = dalloc(ctx, &ctx->rmid, rm.size()))) return rc;
        if ((rc = dalloc(ctx, &ctx->rmid32, rm.size()))) return rc;
        HIPCK(hipMemcpy(ctx->rmid, rm.data(), sizeof(double) * rm.size(), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->rmid32, rm32.data(), sizeof(float) * rm.size(), hipMemcpyHostToDevice));
    } else {
        dalloc(ctx, &ctx->rmid, 0);
        dalloc(ctx, &ctx->rmid32, 0);
    }
    // vectors
    const size_t N = ctx->N;
    int rc;
    double** v64[] = {&ctx->x, &ctx->r, &ctx->p, &ctx->p2, &ctx->qv, &ctx->z, &ctx->dinv, &ctx->F, &ctx->tmp,
                      &ctx->la, &ctx->lw, &ctx->lcp};
    for (auto pp : v64)
        if ((rc = dalloc(ctx, pp, N))) return rc;
    float** v32[] = {&ctx->x32, &ctx->r32, &ctx->p32, &ctx->p2_32, &ctx->q32, &ctx->z32, &ctx->dinv32,
                     &ctx->la32, &ctx->lw32, &ctx->lcp32};
    for (auto pp : v32)
        if ((rc = dalloc(ctx, pp, N))) return rc;
    if ((rc = dalloc(ctx, &ctx->kh, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->kv, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->q, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->kh32, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->kv32, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->dmask, (N + 63) / 64))) return rc;
    if ((rc = dalloc(ctx, &ctx->xprev, N))) return rc;
    if ((rc = dalloc(ctx, &ctx->cur, ctx->E * dim))) return rc;
    if ((rc = dalloc(ctx, &ctx->heat, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->emean, ctx->E))) return rc;
    if ((rc = dalloc(ctx, &ctx->dout, 8))) return rc;
    HIPCK(hipMemset(ctx->cur, 0, sizeof(double) * ctx->E * dim));
    HIPCK(hipMemset(ctx->xprev, 0, sizeof(double) * N));
    ctx->have_law = ctx->have_sel = ctx->have_noheat = false;
    ctx->have_cur = ctx->have_flux = false;
    dalloc(ctx, &ctx->flux, 0);
    dalloc(ctx, &ctx->kpost, 0);
    ctx->nje = 0;
    ctx->dyn = false;  // a transient operator and a boundary matrix refer to the old grid
    ctx->dsh_on = ctx->have_dconv = ctx->have_bnd = false;
    ctx->nbr = 0;
    for (double** pp : {&ctx->dsh, &ctx->dcap, &ctx->F0, &ctx->khB, &ctx->kvB, &ctx->ydir, &ctx->dconv, &ctx->bnd_vals})
        dalloc(ctx, pp, 0);
    dalloc(ctx, &ctx->dsh32, 0);
    dalloc(ctx, &ctx->bnd_vals32, 0);
    dalloc(ctx, &ctx->bnd_rows, 0);
    dalloc(ctx, &ctx->bnd_cols, 0);
    dalloc(ctx, &ctx->bnd_ptr, 0);
    dalloc(ctx, &ctx->zmask, 0);
    ctx->ydir_cap = 0;
    HIPCK(hipMemset(ctx->dmask, 0, sizeof(uint64_t) * ((N + 63) / 64)));
    HIPCK(hipMemset(ctx->q, 0, sizeof(double) * ctx->E));
    HIPCK(hipMemset(ctx->x, 0, sizeof(double) * N));
    ctx->nd = 0;
    ctx->nl = 0;  // nodal loads refer to the old grid
    ctx->dvals_sq = 0.;
    ctx->have_coeffs = false;
    ctx->have_source = false;
    ctx->have_solution = false;
    // launch shapes: tile blocks of BY node rows (BY - 1 owned) x whole x rows, KC planes;
    // measured best at 128 x 128 x 256: 4 rows, 16 planes in both precisions (FP64: 8.23 ms
    // per multigrid solve against 8.32 with 8 planes, 8.47 with 32, 8.32 with 6 rows; a
    // prefetch depth of 2 planes is slower, 8.52)
    if (dim == 3) {
        const int by = (int)std::max<int64_t>(2, std::min<int64_t>(16, 512 / n0));
        ctx->BY[0] = ctx->BY[1] = by;
        ctx->KC[0] = (int)std::max<int64_t>(1, std::min<int64_t>(16, n2));
        ctx->KC[1] = (int)std::max<int64_t>(1, std::min<int64_t>(16, n2));
        // long x rows: the tile kernel's LDS would pass 64 KiB, use row blocks
        const int kind = tile_lds(by, (int)n0) > 65536 ? 1 : 0;
        ctx->apply_kind[0] = ctx->apply_kind[1] = kind;
    } else {
        ctx->KC2 = (int)std::max<int64_t>(4, std::min<int64_t>(64, n1 / 64));
    }
    ctx->upd_blocks = grid1(N, 256, 1024);
    size_t napply = 0;
    if (dim == 3) {
        for (int pk = 0; pk < 2; ++pk) {
            int ytiles = (int)((n1 + ctx->BY[pk] - 2) / (ctx->BY[pk] - 1));
            napply = std::max(napply, (size_t)ytiles * ((n2 + ctx->KC[pk] - 1) / ctx->KC[pk]));
        }
        // row-block apply: one block per row and chunk of KR >= 4 planes, grid rounded to 8
        const size_t nrow = (size_t)n1 * ((n2 + 3) / 4) + 8;
        napply = std::max(napply, nrow);
    } else {
        napply = (size_t)((n1 + ctx->KC2 - 1) / ctx->KC2);
    }
    ctx->npartials = 2 * std::max<size_t>(napply, 4096);
    if ((rc = dalloc(ctx, &ctx->partials, ctx->npartials))) return rc;
    if ((rc = line_config(ctx))) return rc;
    ctx->dirty = true;
    return PF_OK;
}

int pf_coeffs_set(pf_ctx* ctx, int equation, const double* k_inplane, const double* k_vert, int64_t nelem) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (nelem != ctx->E || !k_inplane || !k_vert) return fail(ctx, PF_EINVAL, "coefficient arrays must have one value per element");
    HIPCK(hipSetDevice(ctx->device));
    for (int64_t e = 0; e < nelem; ++e)
        if (!std::isfinite(k_inplane[e]) || !std::isfinite(k_vert[e]) || k_vert[e] <= 0. || k_inplane[e] < 0.)
            return fail(ctx, PF_EINVAL, "conductivities must be finite, in-plane >= 0, vertical > 0 (element " +
                                            std::to_string(e) + ")");
    ctx->equation = equation;
    HIPCK(hipMemcpyAsync(ctx->kh, k_inplane, sizeof(double) * nelem, hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->kv, k_vert, sizeof(double) * nelem, hipMemcpyHostToDevice, ctx->stream));
    int b = grid1(nelem);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(b), dim3(256), 0, ctx->stream, (long long)nelem, ctx->kh, ctx->kh32);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(b), dim3(256), 0, ctx->stream, (long long)nelem, ctx->kv, ctx->kv32);
    dyn_capture(ctx);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->have_coeffs = true;
    ctx->dirty = true;
    return PF_OK;
}

int pf_source_set(pf_ctx* ctx, const double* q_elem, int64_t nelem) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    HIPCK(hipSetDevice(ctx->device));
    if (!q_elem) {
        HIPCK(hipMemsetAsync(ctx->q, 0, sizeof(double) * ctx->E, ctx->stream));
    } else {
        if (nelem != ctx->E) return fail(ctx, PF_EINVAL, "heat array must have one value per element");
        HIPCK(hipMemcpyAsync(ctx->q, q_elem, sizeof(double) * nelem, hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->have_source = q_elem != nullptr;
    ctx->dirty = true;
    return PF_OK;
}

int pf_dirichlet_set(pf_ctx* ctx, const int64_t* nodes, const double* vals, int64_t n) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (n < 0 || (n > 0 && (!nodes || !vals))) return fail(ctx, PF_EINVAL, "bad Dirichlet list");
    HIPCK(hipSetDevice(ctx->device));
    // last value wins for repeated nodes (setBC applied in list order, matrix.hpp:111-118)
    std::vector<uint64_t> mask((ctx->N + 63) / 64, 0ull);
    std::vector<int64_t> pos(0);
    std::vector<long long> un;
    std::vector<double> uv;
    un.reserve(n);
    uv.reserve(n);
    std::vector<int64_t> where;  // node -> index in un (only for repeated detection)
    for (int64_t i = 0; i < n; ++i) {
        int64_t nd = nodes[i];
        if (nd < 0 || nd >= ctx->N) return fail(ctx, PF_EINVAL, "Dirichlet node out of range");
        if (!std::isfinite(vals[i])) return fail(ctx, PF_EINVAL, "Dirichlet value not finite");
        uint64_t bit = 1ull << (nd & 63);
        if (mask[nd >> 6] & bit) {
            for (size_t k = 0; k < un.size(); ++k)
                if (un[k] == nd) uv[k] = vals[i];
            continue;
        }
        mask[nd >> 6] |= bit;
        un.push_back(nd);
        uv.push_back(vals[i]);
    }
    ctx->nd = (int64_t)un.size();
    double sq = 0.;
    for (double v : uv) sq += v * v;
    ctx->dvals_sq = sq;
    int rc;
    if ((rc = dalloc(ctx, &ctx->dnodes, std::max<size_t>(1, un.size())))) return rc;
    if ((rc = dalloc(ctx, &ctx->dvals, std::max<size_t>(1, un.size())))) return rc;
    if (!un.empty()) {
        HIPCK(hipMemcpy(ctx->dnodes, un.data(), sizeof(long long) * un.size(), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->dvals, uv.data(), sizeof(double) * uv.size(), hipMemcpyHostToDevice));
    }
    HIPCK(hipMemcpy(ctx->dmask, mask.data(), sizeof(uint64_t) * mask.size(), hipMemcpyHostToDevice));
    ctx->dirty = true;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_interp_nodes(pf_ctx* ctx, const double* pts, int64_t npts, double fill, double* out) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (npts < 0 || (npts > 0 && (!pts || !out))) return fail(ctx, PF_EINVAL, "bad point list");
    if (npts == 0) return PF_OK;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no solution on the device");
    HIPCK(hipSetDevice(ctx->device));
    const int dim = ctx->dim;
    const size_t nax = ctx->ax[0].size() + ctx->ax[1].size() + (dim == 3 ? ctx->ax[2].size() : 0);
    std::vector<double> axcat;
    axcat.reserve(nax);
    for (int a = 0; a < dim; ++a) axcat.insert(axcat.end(), ctx->ax[a].begin(), ctx->ax[a].end());
    int rc;
    if ((size_t)npts > ctx->ipts_cap) {
        if ((rc = dalloc(ctx, &ctx->ipts, (size_t)npts * dim))) return rc;
        ctx->ipts_cap = (size_t)npts;
    }
    if ((size_t)npts > ctx->iout_cap) {
        if ((rc = dalloc(ctx, &ctx->iout, (size_t)npts))) return rc;
        ctx->iout_cap = (size_t)npts;
    }
    if ((rc = dalloc(ctx, &ctx->iax, nax))) return rc;
    HIPCK(hipMemcpyAsync(ctx->iax, axcat.data(), sizeof(double) * nax, hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->ipts, pts, sizeof(double) * (size_t)npts * dim, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_interp_nodes, dim3(grid1(npts, 256, 8192)), dim3(256), 0, ctx->stream, (const double*)ctx->iax,
                       dim, (int)ctx->n[0], (int)ctx->n[1], dim == 3 ? (int)ctx->n[2] : 1, (const double*)ctx->x,
                       (long long)npts, (const double*)ctx->ipts, fill, ctx->iout);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, ctx->iout, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_load_set(pf_ctx* ctx, const int64_t* nodes, const double* vals, int64_t n) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (n < 0 || (n > 0 && (!nodes || !vals))) return fail(ctx, PF_EINVAL, "bad load list");
    if (n == 0 && ctx->nl == 0) return PF_OK;
    HIPCK(hipSetDevice(ctx->device));
    // one entry per node, repeated nodes summed in list order
    std::vector<std::pair<long long, int64_t>> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (nodes[i] < 0 || nodes[i] >= ctx->N) return fail(ctx, PF_EINVAL, "load node out of range");
        if (!std::isfinite(vals[i])) return fail(ctx, PF_EINVAL, "load value not finite");
        order[(size_t)i] = {(long long)nodes[i], i};
    }
    std::sort(order.begin(), order.end());
    std::vector<long long> un;
    std::vector<double> uv;
    for (const auto& o : order) {
        if (un.empty() || un.back() != o.first) {
            un.push_back(o.first);
            uv.push_back(0.);
        }
        uv.back() += vals[o.second];
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->lnodes, std::max<size_t>(1, un.size())))) return rc;
    if ((rc = dalloc(ctx, &ctx->lvals, std::max<size_t>(1, un.size())))) return rc;
    if (!un.empty()) {
        HIPCK(hipMemcpy(ctx->lnodes, un.data(), sizeof(long long) * un.size(), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->lvals, uv.data(), sizeof(double) * uv.size(), hipMemcpyHostToDevice));
    }
    ctx->nl = (int64_t)un.size();
    ctx->dirty = true;
    return PF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// kernel launch helpers
// ---------------------------------------------------------------------------
template <typename T> static Axes3<T> axes3(pf_ctx* ctx);
// (with the diagonal shift -- capacity and boundary-matrix diagonal -- when one is set)
template <> Axes3<double> axes3<double>(pf_ctx* ctx) {
    return Axes3<double>{(int)ctx->n[0], (int)ctx->n[1], (int)ctx->n[2], ctx->hx, ctx->hy, ctx->hz,
                         ctx->dsh_on ? ctx->dsh : nullptr};
}
template <> Axes3<float> axes3<float>(pf_ctx* ctx) {
    return Axes3<float>{(int)ctx->n[0], (int)ctx->n[1], (int)ctx->n[2], ctx->hx32, ctx->hy32, ctx->hz32,
                        ctx->dsh_on ? ctx->dsh32 : nullptr};
}
template <typename T> static Axes2<T> axes2(pf_ctx* ctx);
template <> Axes2<double> axes2<double>(pf_ctx* ctx) {
    return Axes2<double>{(int)ctx->n[0], (int)ctx->n[1], ctx->hx, ctx->hy, ctx->rmid, ctx->dsh_on ? ctx->dsh : nullptr};
}
template <> Axes2<float> axes2<float>(pf_ctx* ctx) {
    return Axes2<float>{(int)ctx->n[0], (int)ctx->n[1], ctx->hx32, ctx->hy32, ctx->rmid32,
                        ctx->dsh_on ? ctx->dsh32 : nullptr};
}

// off-diagonal boundary-matrix part: MODE_OP y += M Kb M x, MODE_RES y -= M Kb x
template <typename T, int MODE> static void bnd_apply(pf_ctx* ctx, const T* x, T* y, const PcgState* st) {
    const T* vals = (const T*)(sizeof(T) == 8 ? (const void*)ctx->bnd_vals : (const void*)ctx->bnd_vals32);
    hipLaunchKernelGGL((k_bnd_apply<T, MODE>), dim3(grid1(ctx->nbr, 256, 4096)), dim3(256), 0, ctx->stream, ctx->nbr,
                       (const long long*)ctx->bnd_rows, (const int*)ctx->bnd_ptr, (const long long*)ctx->bnd_cols, vals,
                       (const uint64_t*)ctx->dmask, x, y, st);
}

// raw: the operator without the Dirichlet elimination and without the capacity diagonal
// (all nodes free), for the B T^n product of the transient step.  The plain operator and
// residual forms include the boundary matrix; the fused CG forms do not (run_cg switches
// to the unfused step when one is set) and neither does MODE_RESG (multigrid residuals:
// the preconditioner sees the boundary matrix's diagonal only).
template <typename T, int MODE, bool FUSEP, bool DOT>
static void launch_apply(pf_ctx* ctx, const T* xin, const T* zin, const T* pin, T* pout, T* y, const T* F, const T* kh,
                         const T* kv, T* xsol = nullptr, bool raw = false, MgProl<T> pr = MgProl<T>()) {
    const int pk = sizeof(T) == 8 ? 0 : 1;
    const uint64_t* dmask = raw ? ctx->zmask : ctx->dmask;
    Axes3<T> g3 = axes3<T>(ctx);
    Axes2<T> g2 = axes2<T>(ctx);
    if (raw) g3.dsh = g2.dsh = nullptr;
    // MODE_RESP (the multigrid residual with the coarse correction read in) exists in the
    // 3D tile kernel only; mg_vcycle fuses the transfers only where that kernel runs
    if (MODE != MODE_RESP && ctx->dim == 3 && ctx->apply_kind[pk] == 1) {
        const int n0 = (int)ctx->n[0], n1 = (int)ctx->n[1], n2 = (int)ctx->n[2];
        const int KR = ctx->KR[pk];
        const int nb = n1 * ((n2 + KR - 1) / KR);
        const int grid = (nb + 7) / 8 * 8;
        const int threads = (n0 + 63) / 64 * 64;
        const size_t shm = sizeof(T) * (size_t)(4 * (n0 + 1));
        if constexpr (MODE != MODE_RESP)
            pf_launch(ctx, (k_apply3r<T, MODE, FUSEP, DOT>), dim3(grid), dim3(threads), shm, ctx->stream,
                               g3, KR, nb, xin, zin, pin, pout, xsol, y, F, dmask, kh, kv, ctx->st,
                               ctx->partials);
    } else if (MODE == MODE_RESP || ctx->dim == 3) {
        const int BX = (int)ctx->n[0], BY = ctx->BY[pk], KC = ctx->KC[pk];
        const int ytiles = (int)((ctx->n[1] + BY - 2) / (BY - 1));
        const int zt = (int)((ctx->n[2] + KC - 1) / KC);
        const int nb = ytiles * zt, grid = (nb + 7) / 8 * 8;
        const size_t shm = tile_lds(BY, BX) / sizeof(double) * sizeof(T);
        // the multigrid residuals on the solve grid in the XCD-aligned order as well
        // (pf_set_multigrid_xcd(2))
        const int xo = (ctx->mg_xcd == 2 && (MODE == MODE_RESG || MODE == MODE_RESP) && ctx->xcd_order) ? 2 : ctx->xcd_order;
        if (ctx->pfd == 2)
            pf_launch(ctx, (k_apply3d<T, MODE, FUSEP, DOT, 2>), dim3(grid), dim3(BX * BY), shm, ctx->stream,
                               g3, BY, KC, ytiles, nb, xo, xin, zin, pin, pout, xsol, y, F, dmask, kh, kv, ctx->st,
                               ctx->partials, pr);
        else
            pf_launch(ctx, (k_apply3d<T, MODE, FUSEP, DOT, 1>), dim3(grid), dim3(BX * BY), shm, ctx->stream,
                               g3, BY, KC, ytiles, nb, xo, xin, zin, pin, pout, xsol, y, F, dmask, kh, kv, ctx->st,
                               ctx->partials, pr);
    } else {
        const int BX = (int)ctx->n[0];
        int zt = (int)((ctx->n[1] + ctx->KC2 - 1) / ctx->KC2);
        size_t shm = sizeof(T) * (size_t)(4 * (BX + 1));
        if constexpr (MODE != MODE_RESP)
            pf_launch(ctx, (k_apply2d<T, MODE, FUSEP, DOT>), dim3(zt), dim3(BX), shm, ctx->stream, g2,
                               ctx->KC2, xin, zin, pin, pout, xsol, y, F, dmask, kh, kv, ctx->st, ctx->partials);
    }
    if constexpr (!FUSEP && !DOT && (MODE == MODE_OP || MODE == MODE_RES)) {
        if (ctx->have_bnd && !raw) bnd_apply<T, MODE>(ctx, xin, y, nullptr);
    }
}

static int mg_refresh(pf_ctx* ctx);

__global__ void k_copy_dev(long long n, const double* __restrict__ a, double* __restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// ---------------------------------------------------------------------------
// Transient (theta-method) operator of DynamicThermalFem{2D,3D}Solver
// (solvers/thermal/dynamic/femT3d.cpp:140-279, femT2d.cpp:140-330):
//   A = theta K + C,  B = -(1 - theta) K + C,  A T^{n+1} = B T^n + F,
// C the lumped heat-capacity matrix (each node of an element gets the element's c,
// the capacity over the time step in the units of K).  The Dirichlet elimination is
// applied to A and F only (femT3d.cpp:250), so a Dirichlet node's new value is its
// boundary value plus (B T^n) of its row.
// ---------------------------------------------------------------------------

// node capacity: sum of the capacities of the elements touching the node
__global__ void k_node_capacity(ElemGrid eg, const double* __restrict__ ce, double* __restrict__ dcap) {
    const long long N = (long long)eg.n0 * eg.n1 * (eg.dim == 3 ? eg.n2 : 1);
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < N; n += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(n % eg.n0), j = (int)((n / eg.n0) % eg.n1);
        const int k = eg.dim == 3 ? (int)(n / ((long long)eg.n0 * eg.n1)) : 0;
        double s = 0.;
        for (int c = 0; c < (eg.dim == 3 ? 8 : 4); ++c) {
            const int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1), ek = eg.dim == 3 ? k - 1 + (c >> 2) : 0;
            if (ei < 0 || ej < 0 || ek < 0 || ei >= eg.n0 - 1 || ej >= eg.n1 - 1 || (eg.dim == 3 && ek >= eg.n2 - 1))
                continue;
            s += ce[((long long)ek * (eg.n1 - 1) + ej) * (eg.n0 - 1) + ei];
        }
        dcap[n] = s;
    }
}

// kh = theta khB, kv = theta kvB (and the FP32 copies)
__global__ void k_theta_scale(long long E, double theta, const double* __restrict__ khB, const double* __restrict__ kvB,
                              double* __restrict__ kh, double* __restrict__ kv, float* __restrict__ kh32,
                              float* __restrict__ kv32) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        const double a = theta * khB[e], b = theta * kvB[e];
        kh[e] = a;
        kv[e] = b;
        kh32[e] = (float)a;
        kv32[e] = (float)b;
    }
}

// F = F0 + C T - (1 - theta) K T  (= B T^n + F with KT = K T over all nodes); the
// Dirichlet rows' B T^n saved in ydir
__global__ void k_dyn_rhs(long long N, double omt, const double* __restrict__ F0, const double* __restrict__ dsh,
                          const double* __restrict__ T, const double* __restrict__ KT, double* __restrict__ F) {
    for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < N; n += (long long)gridDim.x * blockDim.x)
        F[n] = F0[n] + (dsh[n] * T[n] - omt * KT[n]);
}

__global__ void k_dyn_dir(long long nd, double omt, const long long* __restrict__ dn, const double* __restrict__ dsh,
                          const double* __restrict__ T, const double* __restrict__ KT, double* __restrict__ ydir) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nd; i += (long long)gridDim.x * blockDim.x) {
        const long long n = dn[i];
        ydir[i] = dsh[n] * T[n] - omt * KT[n];
    }
}

__global__ void k_dyn_dir_add(long long nd, const long long* __restrict__ dn, const double* __restrict__ ydir,
                              double* __restrict__ T) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nd; i += (long long)gridDim.x * blockDim.x)
        T[dn[i]] += ydir[i];
}

// theta K into kh/kv from the stiffness as set (khB/kvB)
static void dyn_apply(pf_ctx* ctx) {
    hipLaunchKernelGGL(k_theta_scale, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, (long long)ctx->E,
                       ctx->theta, (const double*)ctx->khB, (const double*)ctx->kvB, ctx->kh, ctx->kv, ctx->kh32,
                       ctx->kv32);
}

// after kh/kv were (re)set to the stiffness: keep it in khB/kvB, scale kh/kv by theta
static void dyn_capture(pf_ctx* ctx) {
    if (!ctx->dyn) return;
    hipLaunchKernelGGL(k_copy_dev, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, (long long)ctx->E,
                       (const double*)ctx->kh, ctx->khB);
    hipLaunchKernelGGL(k_copy_dev, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, (long long)ctx->E,
                       (const double*)ctx->kv, ctx->kvB);
    dyn_apply(ctx);
}

// Partitioned line factors of one grid from the line coefficients in ctx->sb / ctx->sc
// (k_line_bc3d / k_line_bc2d): segments in parallel, then the separator recurrence per
// column.
template <typename T>
static void part_factor(pf_ctx* ctx, int ncol, int n, int m, T* la, T* lw, T* lcp, T* schur) {
    const long long nseg = (long long)((n + m - 1) / m) * ncol;
    hipLaunchKernelGGL((k_part_factor<T>), dim3(grid1(nseg, 256, 1 << 20)), dim3(256), 0, ctx->stream, ncol, n, m,
                       (long long)ncol, (const double*)ctx->sb, (const double*)ctx->sc, la, lw, lcp, ctx->sff, ctx->sfl);
    hipLaunchKernelGGL((k_part_schur<T>), dim3(grid1(ncol, 64, 1 << 20)), dim3(64), 0, ctx->stream, ncol, n, m,
                       (long long)ncol, (const T*)la, (const T*)lw, (const T*)lcp, (const double*)ctx->sff,
                       (const double*)ctx->sfl, schur);
}

// F, Dinv (FP64/FP32) and ||B||: B = M (F - A0 xD) + xD, xD = Dirichlet values
static int prepare(pf_ctx* ctx) {
    if (!ctx->dirty) return PF_OK;
    if (!ctx->have_coeffs) return fail(ctx, PF_ESTATE, "pf_coeffs_set first");
    const long long N = ctx->N;
    const int gb = grid1(N, 256, 4096);
    if (ctx->dim == 3) {
        hipLaunchKernelGGL(k_load3d, dim3(gb), dim3(256), 0, ctx->stream, axes3<double>(ctx), ctx->q, ctx->F);
        hipLaunchKernelGGL((k_diag3d<double>), dim3(gb), dim3(256), 0, ctx->stream, axes3<double>(ctx), ctx->kh, ctx->kv,
                           ctx->dmask, ctx->dinv);
        hipLaunchKernelGGL((k_diag3d<float>), dim3(gb), dim3(256), 0, ctx->stream, axes3<double>(ctx), ctx->kh, ctx->kv,
                           ctx->dmask, ctx->dinv32);
        const int cb = grid1(ctx->n[0] * ctx->n[1], 64, 1 << 20);
        const int ncol = (int)(ctx->n[0] * ctx->n[1]), n2 = (int)ctx->n[2];
        if (ctx->pm64 || ctx->pm32)
            hipLaunchKernelGGL(k_line_bc3d, dim3(gb), dim3(256), 0, ctx->stream, axes3<double>(ctx), ctx->kh, ctx->kv,
                               ctx->dmask, ctx->sb, ctx->sc);
        if (ctx->pm64)
            part_factor<double>(ctx, ncol, n2, ctx->pm64, ctx->la, ctx->lw, ctx->lcp, ctx->schur);
        else
            hipLaunchKernelGGL((k_line_setup3d<double>), dim3(cb), dim3(64), 0, ctx->stream, axes3<double>(ctx),
                               ctx->kh, ctx->kv, ctx->dmask, ctx->la, ctx->lw, ctx->lcp);
        if (ctx->pm32)
            part_factor<float>(ctx, ncol, n2, ctx->pm32, ctx->la32, ctx->lw32, ctx->lcp32, ctx->schur32);
        else
            hipLaunchKernelGGL((k_line_setup3d<float>), dim3(cb), dim3(64), 0, ctx->stream, axes3<double>(ctx),
                               ctx->kh, ctx->kv, ctx->dmask, ctx->la32, ctx->lw32, ctx->lcp32);
    } else {
        hipLaunchKernelGGL(k_load2d, dim3(gb), dim3(256), 0, ctx->stream, axes2<double>(ctx), ctx->q, ctx->F);
        hipLaunchKernelGGL((k_diag2d<double>), dim3(gb), dim3(256), 0, ctx->stream, axes2<double>(ctx), ctx->kh, ctx->kv,
                           ctx->dmask, ctx->dinv);
        hipLaunchKernelGGL((k_diag2d<float>), dim3(gb), dim3(256), 0, ctx->stream, axes2<double>(ctx), ctx->kh, ctx->kv,
                           ctx->dmask, ctx->dinv32);
        const int cb = grid1(ctx->n[0], 64, 1 << 20);
        const int ncol = (int)ctx->n[0], n1 = (int)ctx->n[1];
        if (ctx->pm64 || ctx->pm32)
            hipLaunchKernelGGL(k_line_bc2d, dim3(gb), dim3(256), 0, ctx->stream, axes2<double>(ctx), ctx->kh, ctx->kv,
                               ctx->dmask, ctx->sb, ctx->sc);
        if (ctx->pm64)
            part_factor<double>(ctx, ncol, n1, ctx->pm64, ctx->la, ctx->lw, ctx->lcp, ctx->schur);
        else
            hipLaunchKernelGGL((k_line_setup2d<double>), dim3(cb), dim3(64), 0, ctx->stream, axes2<double>(ctx),
                               ctx->kh, ctx->kv, ctx->dmask, ctx->la, ctx->lw, ctx->lcp);
        if (ctx->pm32)
            part_factor<float>(ctx, ncol, n1, ctx->pm32, ctx->la32, ctx->lw32, ctx->lcp32, ctx->schur32);
        else
            hipLaunchKernelGGL((k_line_setup2d<float>), dim3(cb), dim3(64), 0, ctx->stream, axes2<double>(ctx),
                               ctx->kh, ctx->kv, ctx->dmask, ctx->la32, ctx->lw32, ctx->lcp32);
    }
    if (ctx->equation == PF_ELECTRICAL || !ctx->have_source) {
        // electrical: no load vector (electr2d.cpp:268-317); thermal without source: zero
        hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->F, 0.);
    }
    if (ctx->nl)
        hipLaunchKernelGGL(k_scatter_add, dim3(grid1(ctx->nl)), dim3(256), 0, ctx->stream, (long long)ctx->nl,
                           ctx->lnodes, ctx->lvals, ctx->F);
    // xD into tmp, then r = M (F - A0 xD)
    hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->tmp, 0.);
    if (ctx->nd)
        hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd, ctx->dnodes,
                           ctx->dvals, ctx->tmp);
    if (ctx->dyn)  // the transient step adds B T^n to this static load every step
        hipLaunchKernelGGL(k_copy_dev, dim3(gb), dim3(256), 0, ctx->stream, N, (const double*)ctx->F, ctx->F0);
    launch_apply<double, MODE_RES, false, false>(ctx, ctx->tmp, nullptr, nullptr, nullptr, ctx->r, ctx->F, ctx->kh, ctx->kv);
    hipLaunchKernelGGL((k_norm2<double>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, ctx->r, ctx->st,
                       ctx->partials, 0);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->bnorm = std::sqrt(ctx->hst->bb);  // free rows
    if (int rc = mg_refresh(ctx)) return rc;
    ctx->dirty = false;
    return PF_OK;
}

template <typename T, bool START>
static void launch_part(pf_ctx* ctx, const PartShape& ps, int ncol, int n, long long plane, T* r, const T* q,
                        const T* la, const T* lw, const T* lcp, const T* schur, T* z) {
    const dim3 grid((ncol + ps.CB - 1) / ps.CB);
    const size_t lds = sizeof(T) * (size_t)part_lds_values(ps.P, ps.CB, ps.threads);
#define PF_PART(MM)                                                                                            \
    pf_launch(ctx, (k_cg_part<T, MM, START>), grid, dim3(ps.threads), lds, ctx->stream, ncol, n, ps.P, ps.CB, \
                       plane, r, q, la, lw, lcp, schur, z, ctx->st, ctx->partials)
    switch (ps.m) {
        case 4: PF_PART(4); break;
        case 8: PF_PART(8); break;
        case 16: PF_PART(16); break;
        case 32: PF_PART(32); break;
    }
#undef PF_PART
}

// ---------------------------------------------------------------------------
// Semicoarsening multigrid preconditioner.  Level 0 is the solve grid; each coarser
// level keeps every other node of the in-plane axes (3D: axes 0 and 1; 2D: axis 0) and
// the last one, and every plane of the layer axis.  Coarse operators are the same Q1 operator on the
// coarse grid with area-weighted mean conductivities; the smoother is the damped
// exact line solve along the layer axis (the line preconditioner of the single-level
// PCG), one sweep before and one after the coarse correction, and
// mg_coarse_sweeps sweeps on the coarsest level.  Restriction is the transpose of the
// Q1 interpolation, so the V-cycle is a symmetric positive definite preconditioner
// when omega * lambda_max(M_line^-1 A) < 2.  Element by element, the column-block part
// of a Q1 box matrix with diagonal conductivity is kx'/3 + ky'/3 (x) M_z + kz'/9 S_z and
// the ratio of the two matrices' eigenvalues is at most 3 (reached for kx' >> ky', kz'),
// so lambda_max <= 3 for any element shapes: omega <= 0.66 is accepted, 0.65 default.
// ---------------------------------------------------------------------------

// planes along the layer (line) axis: axis 2 in 3D, axis 1 in 2D
static int mg_planes(const pf_ctx* ctx) { return (int)(ctx->dim == 3 ? ctx->n[2] : ctx->n[1]); }

// Capacity diagonal of a coarse level: c_c = P^T c_f, the row sums of P^T C P (P
// reproduces constants), i.e. the lumped form of the Galerkin coarse capacity matrix.
__global__ void k_restrict_diag(MgTransfer tr, const double* __restrict__ wx, const double* __restrict__ wy,
                                const double* __restrict__ cf, double* __restrict__ cc, float* __restrict__ cc32) {
    const long long planec = (long long)tr.n0c * tr.n1c, planef = (long long)tr.n0f * tr.n1f;
    const long long Nc = planec * tr.n2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nc;
         idx += (long long)gridDim.x * blockDim.x) {
        const int I = (int)(idx % tr.n0c), J = (int)((idx / tr.n0c) % tr.n1c);
        const long long k = idx / planec;
        int ii[3], jj[3];
        double wi[3], wj[3];
        restrict_axis(I, tr.n0c, tr.cx, tr.mx, wx, ii, wi);
        restrict_axis(J, tr.n1c, tr.cy, tr.my, wy, jj, wj);
        const double* src = cf + k * planef;
        double s = 0.;
        for (int b = 0; b < 3; ++b) {
            const double* row = src + (long long)jj[b] * tr.n0f;
            s += wj[b] * (wi[0] * row[ii[0]] + wi[1] * row[ii[1]] + wi[2] * row[ii[2]]);
        }
        cc[idx] = s;
        cc32[idx] = (float)s;
    }
}

static MgTransfer mg_transfer(pf_ctx* ctx, size_t l) {
    const MgLevel& L = ctx->mg[l];
    MgTransfer tr;
    tr.n0f = l ? ctx->mg[l - 1].n0 : (int)ctx->n[0];
    tr.n1f = l ? ctx->mg[l - 1].n1 : (ctx->dim == 3 ? (int)ctx->n[1] : 1);
    tr.n0c = L.n0;
    tr.n1c = L.n1;
    tr.n2 = mg_planes(ctx);
    tr.cx = L.cx;
    tr.cy = L.cy;
    tr.mx = L.mx;
    tr.rx = L.rx;
    tr.ry = L.ry;
    tr.my = L.my;
    return tr;
}

// coarse element data, Dirichlet masks and line factors of every level (after a
// change of the coefficients or of the Dirichlet set)
static int mg_refresh(pf_ctx* ctx) {
    if (!ctx->mg_built) return PF_OK;
    if (!ctx->pm64 || !ctx->pm32) {  // sequential line solve selected: levels dropped
        mg_free(ctx);
        return PF_OK;
    }
    const int n2 = mg_planes(ctx);
    int rc;
    for (size_t l = 0; l < ctx->mg.size(); ++l) {
        MgLevel& L = ctx->mg[l];
        const MgLevel* F = l ? &ctx->mg[l - 1] : nullptr;
        const MgTransfer tr = mg_transfer(ctx, l);
        if (ctx->dim == 3)
            hipLaunchKernelGGL(k_mg_coarsen_k, dim3(grid1(L.E, 256, 4096)), dim3(256), 0, ctx->stream, tr,
                               F ? F->hx : ctx->hx, F ? F->hy : ctx->hy, F ? F->kh : ctx->kh, F ? F->kv : ctx->kv,
                               L.kh, L.kv);
        else
            hipLaunchKernelGGL(k_mg_coarsen_k2d, dim3(grid1(L.E, 256, 4096)), dim3(256), 0, ctx->stream, tr,
                               F ? F->hx : ctx->hx, F ? F->rmid : ctx->rmid, F ? F->kh : ctx->kh,
                               F ? F->kv : ctx->kv, L.kh, L.kv);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid1(L.E, 256, 4096)), dim3(256), 0, ctx->stream, L.E,
                           L.kh, L.kh32);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid1(L.E, 256, 4096)), dim3(256), 0, ctx->stream, L.E,
                           L.kv, L.kv32);
        hipLaunchKernelGGL(k_mg_coarsen_mask, dim3(grid1((L.N + 63) / 64, 256, 4096)), dim3(256), 0, ctx->stream, tr,
                           F ? F->dmask : ctx->dmask, L.dmask);
        const int ncol = L.n0 * L.n1;
        // Segment lengths: with the default lengths (8 planes FP64, 16 FP32 on the solve
        // grid) the coarse levels, which are latency-bound, take shorter segments, i.e.
        // shorter dependent chains per thread, as long as the separators (P - 1) stay few
        // enough for the scan of the separator system (at most 64): 8 planes, and 4 on
        // levels of at most 512 columns (measured per level on the 4.2M-DOF stack; longer
        // segments were slower on every coarse level).  A requested length applies to
        // every level.
        int m64 = ctx->pm64, m32 = ctx->pm32;
        if (ctx->line_m == 0) {
            int mc = 0;
            if ((n2 + 7) / 8 <= 65) mc = 8;
            if (ncol <= 512 && (n2 + 3) / 4 <= 65) mc = 4;
            if (mc && (size_t)((n2 + mc - 1) / mc) * ncol <= ctx->nseg_cap && part_shape<double>(mc, n2, ncol).m &&
                part_shape<float>(mc, n2, ncol).m) {
                m64 = std::min(m64, mc);
                m32 = std::min(m32, mc);
            }
        }
        const PartShape s64 = part_shape<double>(m64, n2, ncol), s32 = part_shape<float>(m32, n2, ncol);
        if (!s64.m || !s32.m) return fail(ctx, PF_EINVAL, "multigrid: line partition does not fit a coarse level");
        L.pm64 = s64.m;
        L.pm32 = s32.m;
        if (L.schur_m64 != s64.m) {
            if ((rc = dalloc(ctx, &L.schur, 3 * (size_t)std::max(s64.P - 1, 1) * ncol))) return rc;
            L.schur_m64 = s64.m;
        }
        if (L.schur_m32 != s32.m) {
            if ((rc = dalloc(ctx, &L.schur32, 3 * (size_t)std::max(s32.P - 1, 1) * ncol))) return rc;
            L.schur_m32 = s32.m;
        }
        const int gl = grid1((long long)ncol * n2, 256, 4096);
        if (ctx->dsh_on) {
            if (!L.dsh && ((rc = dalloc(ctx, &L.dsh, L.N)) || (rc = dalloc(ctx, &L.dsh32, L.N)))) return rc;
            hipLaunchKernelGGL(k_restrict_diag, dim3(grid1(L.N, 256, 4096)), dim3(256), 0, ctx->stream, tr, L.wx,
                               L.wy, (const double*)(F ? F->dsh : ctx->dsh), L.dsh, L.dsh32);
        }
        if (ctx->dim == 3) {
            const Axes3<double> g{L.n0, L.n1, n2, L.hx, L.hy, ctx->hz, ctx->dsh_on ? L.dsh : nullptr};
            hipLaunchKernelGGL(k_line_bc3d, dim3(gl), dim3(256), 0, ctx->stream, g, L.kh, L.kv, L.dmask, ctx->sb,
                               ctx->sc);
        } else {
            const Axes2<double> g{L.n0, n2, L.hx, ctx->hy, L.rmid, ctx->dsh_on ? L.dsh : nullptr};
            hipLaunchKernelGGL(k_line_bc2d, dim3(gl), dim3(256), 0, ctx->stream, g, L.kh, L.kv, L.dmask, ctx->sb,
                               ctx->sc);
        }
        part_factor<double>(ctx, ncol, n2, s64.m, L.la, L.lw, L.lcp, L.schur);
        part_factor<float>(ctx, ncol, n2, s32.m, L.la32, L.lw32, L.lcp32, L.schur32);
    }
    HIPCK(hipGetLastError());
    return PF_OK;
}

template <typename T> static int upload(pf_ctx* ctx, T** dst, const std::vector<T>& v) {
    int rc;
    if ((rc = dalloc(ctx, dst, v.size()))) return rc;
    if (v.empty()) return PF_OK;
    HIPCK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return PF_OK;
}

// coarse index list, per-fine-node interval and weight, coarse coordinates of one axis
static void mg_coarsen_axis(const std::vector<double>& f, bool on, std::vector<int>& ci, std::vector<int>& m,
                            std::vector<double>& w, std::vector<double>& c) {
    const int n = (int)f.size();
    ci.clear();
    for (int i = 0; i < n; i += on ? 2 : 1) ci.push_back(i);
    if (ci.back() != n - 1) ci.push_back(n - 1);
    m.assign(n, 0);
    w.assign(n, 0.);
    for (size_t I = 0; I + 1 < ci.size(); ++I)
        for (int i = ci[I]; i < ci[I + 1]; ++i) {
            m[i] = (int)I;
            w[i] = (f[i] - f[ci[I]]) / (f[ci[I + 1]] - f[ci[I]]);
        }
    m[n - 1] = (int)ci.size() - 1;
    w[n - 1] = 0.;
    c.clear();
    for (int i : ci) c.push_back(f[i]);
}

// Restriction rows of one axis, per coarse node I: its (up to) 3 fine nodes (those strictly
// between the neighbouring coarse nodes; absent ones repeat the first with weight 0) and
// their interpolation weights in I, as restrict_axis computes them on the device, the FP32
// weights in FP32 arithmetic like the kernel's.  Read by k_mg_restrict, so that its index
// and weight loads do not wait on the coarse-index loads.
static void mg_restrict_rows(const std::vector<int>& ci, const std::vector<int>& m, const std::vector<double>& w,
                             std::vector<int>& ri, std::vector<double>& rw, std::vector<float>& rw32) {
    const int nc = (int)ci.size();
    ri.assign(3 * (size_t)nc, 0);
    rw.assign(3 * (size_t)nc, 0.);
    rw32.assign(3 * (size_t)nc, 0.f);
    for (int I = 0; I < nc; ++I) {
        const int lo = I > 0 ? ci[I - 1] + 1 : 0;
        const int hi = I < nc - 1 ? ci[I + 1] - 1 : ci[I];
        for (int a = 0; a < 3; ++a) {
            const int i = lo + a, ix = i <= hi ? i : lo;
            ri[3 * I + a] = ix;
            if (i > hi) continue;
            if (m[ix] == I) {
                rw[3 * I + a] = 1. - w[ix];
                rw32[3 * I + a] = 1.f - (float)w[ix];
            } else if (m[ix] + 1 == I) {
                rw[3 * I + a] = w[ix];
                rw32[3 * I + a] = (float)w[ix];
            }
        }
    }
}

// levels of the current grid: an in-plane axis (3D: 0 and 1, 2D: 0) is coarsened while
// it has more than 3 nodes; at most mg_max_levels levels in all (0: no limit)
static int mg_build(pf_ctx* ctx) {
    if (!ctx->pm64 || !ctx->pm32)
        return fail(ctx, PF_EINVAL, "the multigrid preconditioner needs the partitioned line solve");
    mg_free(ctx);
    int rc;
    const bool three = ctx->dim == 3;
    const int n2 = mg_planes(ctx);
    // 2D: axis 1 is the layer axis; the in-plane y axis of the tables is a single node
    std::vector<double> fx = ctx->ax[0], fy = three ? ctx->ax[1] : std::vector<double>{0.};
    while (ctx->mg_max_levels == 0 || (int)ctx->mg.size() + 1 < ctx->mg_max_levels) {
        const bool cxa = fx.size() > 3, cya = fy.size() > 3;
        if (!cxa && !cya) break;
        std::vector<int> cix, ciy, mx, my;
        std::vector<double> wx, wy, cxv, cyv;
        mg_coarsen_axis(fx, cxa, cix, mx, wx, cxv);
        mg_coarsen_axis(fy, cya, ciy, my, wy, cyv);
        ctx->mg.emplace_back();
        MgLevel& L = ctx->mg.back();
        L.n0 = (int)cxv.size();
        L.n1 = (int)cyv.size();
        L.N = (long long)L.n0 * L.n1 * n2;
        L.E = (long long)(L.n0 - 1) * (three ? L.n1 - 1 : 1) * (n2 - 1);
        std::vector<double> hx(L.n0 - 1), hy(L.n1 - 1), rm;
        for (int i = 0; i + 1 < L.n0; ++i) hx[i] = cxv[i + 1] - cxv[i];
        for (int j = 0; j + 1 < L.n1; ++j) hy[j] = cyv[j + 1] - cyv[j];
        if (ctx->geom == PF_GEOM_CYLINDRICAL) {
            for (int i = 0; i + 1 < L.n0; ++i) rm.push_back(0.5 * (cxv[i] + cxv[i + 1]));
            if ((rc = upload(ctx, &L.rmid, rm)) || (rc = upload(ctx, &L.rmid32, std::vector<float>(rm.begin(), rm.end()))))
                return rc;
        }
        if ((rc = upload(ctx, &L.cx, cix)) || (rc = upload(ctx, &L.cy, ciy)) || (rc = upload(ctx, &L.mx, mx)) ||
            (rc = upload(ctx, &L.my, my)) || (rc = upload(ctx, &L.wx, wx)) || (rc = upload(ctx, &L.wy, wy)) ||
            (rc = upload(ctx, &L.wx32, std::vector<float>(wx.begin(), wx.end()))) ||
            (rc = upload(ctx, &L.wy32, std::vector<float>(wy.begin(), wy.end()))) || (rc = upload(ctx, &L.hx, hx)) ||
            (rc = upload(ctx, &L.hy, hy)) || (rc = upload(ctx, &L.hx32, std::vector<float>(hx.begin(), hx.end()))) ||
            (rc = upload(ctx, &L.hy32, std::vector<float>(hy.begin(), hy.end()))))
            return rc;
        {
            std::vector<int> rx, ry;
            std::vector<double> rwx, rwy;
            std::vector<float> rwx32, rwy32;
            mg_restrict_rows(cix, mx, wx, rx, rwx, rwx32);
            mg_restrict_rows(ciy, my, wy, ry, rwy, rwy32);
            if ((rc = upload(ctx, &L.rx, rx)) || (rc = upload(ctx, &L.ry, ry)) || (rc = upload(ctx, &L.rwx, rwx)) ||
                (rc = upload(ctx, &L.rwy, rwy)) || (rc = upload(ctx, &L.rwx32, rwx32)) || (rc = upload(ctx, &L.rwy32, rwy32)))
                return rc;
        }
        if ((rc = dalloc(ctx, &L.kh, L.E)) || (rc = dalloc(ctx, &L.kv, L.E)) || (rc = dalloc(ctx, &L.kh32, L.E)) ||
            (rc = dalloc(ctx, &L.kv32, L.E)) || (rc = dalloc(ctx, &L.dmask, (L.N + 63) / 64)) ||
            (rc = dalloc(ctx, &L.la, L.N)) || (rc = dalloc(ctx, &L.lw, L.N)) || (rc = dalloc(ctx, &L.lcp, L.N)) ||
            (rc = dalloc(ctx, &L.la32, L.N)) || (rc = dalloc(ctx, &L.lw32, L.N)) ||
            (rc = dalloc(ctx, &L.lcp32, L.N)) || (rc = dalloc(ctx, &L.b, L.N)) || (rc = dalloc(ctx, &L.x, L.N)) ||
            (rc = dalloc(ctx, &L.r, L.N)) || (rc = dalloc(ctx, &L.b32, L.N)) || (rc = dalloc(ctx, &L.x32, L.N)) ||
            (rc = dalloc(ctx, &L.r32, L.N)))
            return rc;
        fx = cxv;
        fy = cyv;
    }
    if ((rc = dalloc(ctx, &ctx->mg_r0, ctx->N)) || (rc = dalloc(ctx, &ctx->mg_r0_32, ctx->N))) return rc;
    ctx->mg_built = true;
    return mg_refresh(ctx);
}

// one level's operands in precision T (level 0: the solve grid, b = the PCG residual,
// x = the preconditioned residual z)
template <typename T>
struct MgView {
    Axes3<T> g;   // 3D levels
    Axes2<T> g2;  // 2D levels
    int ncol, planes;
    const T *kh, *kv, *la, *lw, *lcp, *schur, *wx, *wy, *rwx, *rwy;
    const uint64_t* dmask;
    T *b, *x, *r;
    int pm;
    int level;
    long long nodes() const { return (long long)ncol * planes; }
    long long elems(int dim) const {
        return dim == 3 ? (long long)(g.n0 - 1) * (g.n1 - 1) * (planes - 1) : (long long)(g2.n0 - 1) * (planes - 1);
    }
};

// FP64 / FP32 members of the context and of a level
template <typename T> struct MgPick;
template <> struct MgPick<double> {
    static int pm(const pf_ctx* c) { return c->pm64; }
    static const double* kh(const pf_ctx* c) { return c->kh; }
    static const double* kv(const pf_ctx* c) { return c->kv; }
    static const double* la(const pf_ctx* c) { return c->la; }
    static const double* lw(const pf_ctx* c) { return c->lw; }
    static const double* lcp(const pf_ctx* c) { return c->lcp; }
    static const double* schur(const pf_ctx* c) { return c->schur; }
    static double* r0(const pf_ctx* c) { return c->mg_r0; }
    static const double* hy(const pf_ctx* c) { return c->hy; }
    static const double* hz(const pf_ctx* c) { return c->hz; }
    static const MgLevel& lv(const pf_ctx* c, int l) { return c->mg[l - 1]; }
    static int pm(const MgLevel& L) { return L.pm64; }
    static const double* hx(const MgLevel& L) { return L.hx; }
    static const double* hyl(const MgLevel& L) { return L.hy; }
    static const double* rmid(const MgLevel& L) { return L.rmid; }
    static const double* kh(const MgLevel& L) { return L.kh; }
    static const double* kv(const MgLevel& L) { return L.kv; }
    static const double* la(const MgLevel& L) { return L.la; }
    static const double* lw(const MgLevel& L) { return L.lw; }
    static const double* lcp(const MgLevel& L) { return L.lcp; }
    static const double* schur(const MgLevel& L) { return L.schur; }
    static const double* wx(const MgLevel& L) { return L.wx; }
    static const double* wy(const MgLevel& L) { return L.wy; }
    static const double* rwx(const MgLevel& L) { return L.rwx; }
    static const double* rwy(const MgLevel& L) { return L.rwy; }
    static double* b(const MgLevel& L) { return L.b; }
    static double* x(const MgLevel& L) { return L.x; }
    static double* r(const MgLevel& L) { return L.r; }
    static const double* dsh(const MgLevel& L) { return L.dsh; }
};
template <> struct MgPick<float> {
    static int pm(const pf_ctx* c) { return c->pm32; }
    static const float* kh(const pf_ctx* c) { return c->kh32; }
    static const float* kv(const pf_ctx* c) { return c->kv32; }
    static const float* la(const pf_ctx* c) { return c->la32; }
    static const float* lw(const pf_ctx* c) { return c->lw32; }
    static const float* lcp(const pf_ctx* c) { return c->lcp32; }
    static const float* schur(const pf_ctx* c) { return c->schur32; }
    static float* r0(const pf_ctx* c) { return c->mg_r0_32; }
    static const float* hy(const pf_ctx* c) { return c->hy32; }
    static const float* hz(const pf_ctx* c) { return c->hz32; }
    static const MgLevel& lv(const pf_ctx* c, int l) { return c->mg[l - 1]; }
    static int pm(const MgLevel& L) { return L.pm32; }
    static const float* hx(const MgLevel& L) { return L.hx32; }
    static const float* hyl(const MgLevel& L) { return L.hy32; }
    static const float* rmid(const MgLevel& L) { return L.rmid32; }
    static const float* kh(const MgLevel& L) { return L.kh32; }
    static const float* kv(const MgLevel& L) { return L.kv32; }
    static const float* la(const MgLevel& L) { return L.la32; }
    static const float* lw(const MgLevel& L) { return L.lw32; }
    static const float* lcp(const MgLevel& L) { return L.lcp32; }
    static const float* schur(const MgLevel& L) { return L.schur32; }
    static const float* wx(const MgLevel& L) { return L.wx32; }
    static const float* wy(const MgLevel& L) { return L.wy32; }
    static const float* rwx(const MgLevel& L) { return L.rwx32; }
    static const float* rwy(const MgLevel& L) { return L.rwy32; }
    static float* b(const MgLevel& L) { return L.b32; }
    static float* x(const MgLevel& L) { return L.x32; }
    static float* r(const MgLevel& L) { return L.r32; }
    static const float* dsh(const MgLevel& L) { return L.dsh32; }
};

template <typename T>
static MgView<T> mg_view(pf_ctx* ctx, int l, T* r0, T* z0) {
    using S = MgPick<T>;
    MgView<T> v{};
    v.pm = S::pm(ctx);
    v.planes = mg_planes(ctx);
    v.level = l;
    if (l == 0) {
        if (ctx->dim == 3) v.g = axes3<T>(ctx);
        else v.g2 = axes2<T>(ctx);
        v.ncol = (int)(ctx->dim == 3 ? ctx->n[0] * ctx->n[1] : ctx->n[0]);
        v.kh = S::kh(ctx);
        v.kv = S::kv(ctx);
        v.la = S::la(ctx);
        v.lw = S::lw(ctx);
        v.lcp = S::lcp(ctx);
        v.schur = S::schur(ctx);
        v.wx = v.wy = v.rwx = v.rwy = nullptr;
        v.dmask = ctx->dmask;
        v.b = r0;
        v.x = z0;
        v.r = S::r0(ctx);
        return v;
    }
    const MgLevel& L = S::lv(ctx, l);
    v.pm = S::pm(L);
    const T* dl = ctx->dsh_on ? S::dsh(L) : nullptr;
    if (ctx->dim == 3) v.g = Axes3<T>{L.n0, L.n1, v.planes, S::hx(L), S::hyl(L), S::hz(ctx), dl};
    else v.g2 = Axes2<T>{L.n0, v.planes, S::hx(L), S::hy(ctx), S::rmid(L), dl};
    v.ncol = L.n0 * L.n1;
    v.kh = S::kh(L);
    v.kv = S::kv(L);
    v.la = S::la(L);
    v.lw = S::lw(L);
    v.lcp = S::lcp(L);
    v.schur = S::schur(L);
    v.wx = S::wx(L);
    v.wy = S::wy(L);
    v.rwx = S::rwx(L);
    v.rwy = S::rwy(L);
    v.dmask = L.dmask;
    v.b = S::b(L);
    v.x = S::x(L);
    v.r = S::r(L);
    return v;
}

// r = b - A x on level l (skipped after the stop)
template <typename T>
static void mg_residual_launch(pf_ctx* ctx, int l, const MgView<T>& v, MgProl<T> pr = MgProl<T>()) {
    if (l == 0) {
        if (pr.xc)
            launch_apply<T, MODE_RESP, false, false>(ctx, v.x, nullptr, nullptr, nullptr, v.r, v.b, v.kh, v.kv,
                                                     nullptr, false, pr);
        else
            launch_apply<T, MODE_RESG, false, false>(ctx, v.x, nullptr, nullptr, nullptr, v.r, v.b, v.kh, v.kv);
        return;
    }
    if (ctx->dim == 2) {
        // one block per chunk of KC rows of the layer axis, one thread per node of axis 0;
        // short chunks on the (latency-bound) coarse levels
        const int n0 = v.g2.n0, n1 = v.g2.n1;
        int KC = 64;
        while (KC > 1 && (n1 + KC - 1) / KC < 256) KC >>= 1;
        pf_launch(ctx, (k_apply2d<T, MODE_RESG, false, false>), dim3((n1 + KC - 1) / KC), dim3(n0),
                           sizeof(T) * 4 * (size_t)(n0 + 1), ctx->stream, v.g2, KC, (const T*)v.x, (const T*)nullptr,
                           (const T*)nullptr, (T*)nullptr, (T*)nullptr, v.r, (const T*)v.b, v.dmask, v.kh, v.kv,
                           ctx->st, ctx->partials);
        return;
    }
    const int n0 = v.g.n0, n1 = v.g.n1, n2 = v.g.n2;
    int BY = std::max(2, std::min(16, 512 / n0));
    while (BY > 2 && (BY * n0 > 1024 || tile_lds(BY, n0) > 65536)) --BY;
    {  // fewer rows where that saves waves per z-chunk (tiles x waves per tile; ties keep
       // the larger tile): 17 x 17 columns take 11 rows (2 tiles of 3 waves), not 16 (2 of 5)
        int best = BY;
        long long cost = -1;
        for (int by = std::min(BY, n1 + 1); by >= 2; --by) {
            const long long c = (long long)((n1 + by - 2) / (by - 1)) * ((by * n0 + 63) / 64);
            if (cost < 0 || c < cost) {
                cost = c;
                best = by;
            }
        }
        BY = best;
    }
    const int ytiles = (n1 + BY - 2) / (BY - 1);
    // coarse levels are latency-bound: shorter z-marches (down to 1 plane per block)
    // until the grid has about 2 blocks per CU (aiming at 4 measured slower: 65 x 65
    // columns 22.0 against 17.7 us per launch)
    int KC = 8;
    while (KC > 1 && (long long)ytiles * ((n2 + KC - 1) / KC) < 512) KC >>= 1;
    const int nb = ytiles * ((n2 + KC - 1) / KC), grid = (nb + 7) / 8 * 8;
    const size_t shm = tile_lds(BY, n0) / sizeof(double) * sizeof(T);
    const int xo = ctx->mg_xcd ? 2 : ctx->xcd_order;
    if (pr.xc)
        pf_launch(ctx, (k_apply3d<T, MODE_RESP, false, false, 1>), dim3(grid), dim3(n0 * BY), shm, ctx->stream,
                           v.g, BY, KC, ytiles, nb, xo, (const T*)v.x, (const T*)nullptr,
                           (const T*)nullptr, (T*)nullptr, (T*)nullptr, v.r, (const T*)v.b, v.dmask, v.kh, v.kv,
                           ctx->st, ctx->partials, pr);
    else
        pf_launch(ctx, (k_apply3d<T, MODE_RESG, false, false, 1>), dim3(grid), dim3(n0 * BY), shm, ctx->stream,
                           v.g, BY, KC, ytiles, nb, xo, (const T*)v.x, (const T*)nullptr,
                           (const T*)nullptr, (T*)nullptr, (T*)nullptr, v.r, (const T*)v.b, v.dmask, v.kh, v.kv,
                           ctx->st, ctx->partials, pr);
}

template <typename T>
static void mg_residual(pf_ctx* ctx, int l, const MgView<T>& v, MgProl<T> pr = MgProl<T>(), long long ncoarse = 0) {
    const double extra = pr.xc ? (double)sizeof(T) * (double)ncoarse : 0.;  // the coarse correction read
    const int hk = prof_kern<T>(ctx, PF_K_RESIDUAL, l,
                                bytes_residual(v.nodes(), v.elems(ctx->dim), (int)sizeof(T)) + extra, !ctx->have_bnd);
    mg_residual_launch<T>(ctx, l, v, pr);
    prof_close(ctx, hk);
}

// one smoothing sweep on a level: xout = xin + omega M_line^-1 r (see k_mg_smooth)
template <typename T, int M>
static void mg_smooth_m(pf_ctx* ctx, dim3 grid, int threads, size_t lds, const PartShape& ps, int ncol, int n,
                        long long plane, bool upd, int fin, T* r, const T* q, const T* xin, T* xout, const T* rp,
                        T omega, const MgView<T>& v, int xfk, const MgXfer<T>& xf, int xcd) {
#define PF_MGS(U, F, X)                                                                                          \
    pf_launch(ctx, (k_mg_smooth<T, M, U, F, X>), grid, dim3(threads), lds, ctx->stream, ncol, n, ps.P, ps.CB, \
                       plane, r, q, xin, xout, rp, omega, v.la, v.lw, v.lcp, v.schur, ctx->st, ctx->partials, xf, xcd)
    if (xfk == 1) {  // coarse levels: first sweep from zero, restriction fused
        PF_MGS(false, 0, 1);
    } else if (xfk == 2) {  // post-smoothing with the prolongation fused
        if (fin == 0) PF_MGS(false, 0, 2);
        else if (fin == 1) PF_MGS(false, 1, 2);
        else PF_MGS(false, 2, 2);
    } else if (upd) {
        if (fin == 0) PF_MGS(true, 0, 0);
        else PF_MGS(true, 1, 0);
    } else {
        if (fin == 0) PF_MGS(false, 0, 0);
        else if (fin == 1) PF_MGS(false, 1, 0);
        else PF_MGS(false, 2, 0);
    }
#undef PF_MGS
}

template <typename T>
static void mg_smooth(pf_ctx* ctx, const MgView<T>& v, bool upd, int fin, T* r, const T* q, const T* xin, T* xout,
                      const T* rp, T omega = T(-1), int xfk = 0, const MgXfer<T>& xf = MgXfer<T>(),
                      double xbytes = 0.) {
    const int ncol = v.ncol, n = v.planes;
    PartShape ps = part_shape<T>(v.pm, n, ncol);
    // small coarse levels: fewer columns per block, so that the few columns spread over
    // at least 64 blocks (larger levels keep full blocks, which measured faster)
    while (ps.CB > 4 && (ncol + ps.CB - 1) / ps.CB < 64) {
        ps.CB >>= 1;
        ps.threads = (ps.CB * ps.P + 63) / 64 * 64;
    }
    // coarse levels with the XCD-aligned order: the grid a multiple of 8 (idle blocks
    // past the last column group)
    const int xcd = (ctx->mg_xcd == 2 || (ctx->mg_xcd && v.level > 0)) ? 1 : 0;
    const int ngroups = (ncol + ps.CB - 1) / ps.CB;
    const dim3 grid(xcd ? (ngroups + 7) / 8 * 8 : ngroups);
    const size_t lds = sizeof(T) * (size_t)part_lds_values(ps.P, ps.CB, ps.threads);
    const T om = omega > T(0) ? omega : T(ctx->mg_omega);
    const int hk = prof_kern<T>(ctx, PF_K_SMOOTH, v.level,
                                bytes_smooth(v.nodes(), ncol, ps.P, (int)sizeof(T), upd, xin != nullptr, fin != 0) +
                                    xbytes);
    switch (ps.m) {
        case 4: mg_smooth_m<T, 4>(ctx, grid, ps.threads, lds, ps, ncol, n, ncol, upd, fin, r, q, xin, xout, rp, om, v, xfk, xf, xcd); break;
        case 8: mg_smooth_m<T, 8>(ctx, grid, ps.threads, lds, ps, ncol, n, ncol, upd, fin, r, q, xin, xout, rp, om, v, xfk, xf, xcd); break;
        case 16: mg_smooth_m<T, 16>(ctx, grid, ps.threads, lds, ps, ncol, n, ncol, upd, fin, r, q, xin, xout, rp, om, v, xfk, xf, xcd); break;
        case 32: mg_smooth_m<T, 32>(ctx, grid, ps.threads, lds, ps, ncol, n, ncol, upd, fin, r, q, xin, xout, rp, om, v, xfk, xf, xcd); break;
    }
    prof_close(ctx, hk);
}

// tables of the transfer between level t and t + 1 with a source vector, for the fused forms
template <typename T>
static MgXfer<T> mg_xfer(pf_ctx* ctx, int t, const MgView<T>& coarse, const T* src, const uint64_t* dmask) {
    const MgTransfer tr = mg_transfer(ctx, (size_t)t);
    MgXfer<T> x;
    x.n0f = tr.n0f;
    x.n1f = tr.n1f;
    x.n0c = tr.n0c;
    x.n1c = tr.n1c;
    x.cx = tr.cx;
    x.cy = tr.cy;
    x.mx = tr.mx;
    x.my = tr.my;
    x.wx = coarse.wx;
    x.wy = coarse.wy;
    x.src = src;
    x.dmask = dmask;
    return x;
}

template <typename T> static MgProl<T> mg_prol(const MgXfer<T>& x) {
    MgProl<T> p;
    p.n0c = x.n0c;
    p.n1c = x.n1c;
    p.mx = x.mx;
    p.my = x.my;
    p.wx = x.wx;
    p.wy = x.wy;
    p.xc = x.src;
    return p;
}

// z = V(r): one V-cycle.  upd: apply r -= alpha q first (inside the first sweep);
// fin: 1 the last sweep ends the PCG iteration, 2 it starts a PCG run.
// 3D with the tile operator kernel: the grid transfers run inside the neighbouring
// kernels (restriction in a coarse level's first sweep, prolongation in the residual and
// the sweep that follow it), which saves two launches and a vector pass per level.
template <typename T>
static void mg_vcycle(pf_ctx* ctx, T* r, const T* q, T* z, bool upd, int fin) {
    const int L = (int)ctx->mg.size() + 1;
    const bool fuse = ctx->mg_fuse && ctx->dim == 3 && L > 1 && ctx->apply_kind[sizeof(T) == 8 ? 0 : 1] == 0;
    // profiled chunks: level-0 kernels by kind, all coarse-level work as one record
    auto cat = [](int l, int kind) { return l == 0 ? kind : -1; };
    auto open = [&](int c) { return c >= 0 ? prof_open(ctx, c) : -1; };
    int hc = -1;
    for (int l = 0; l < L; ++l) {
        const MgView<T> v = mg_view<T>(ctx, l, r, z);
        const bool last = l == L - 1;
        const int sweeps = last ? ctx->mg_coarse_sweeps : 1;
        if (l == 1) hc = prof_open(ctx, PF_KCAT_MG_COARSE);
        for (int s = 0; s < sweeps; ++s) {
            const int f = (l == 0 && last && s == sweeps - 1) ? fin : 0;
            if (s > 0) {
                const int h = open(cat(l, PF_KCAT_MG_RESIDUAL));
                mg_residual<T>(ctx, l, v);
                prof_close(ctx, h);
            }
            const int h = open(cat(l, PF_KCAT_MG_SMOOTH));
            if (s == 0 && fuse && l > 0) {
                // b = R r of the finer level, read and stored by the sweep itself
                const MgView<T> fv = mg_view<T>(ctx, l - 1, r, z);
                const MgXfer<T> xf = mg_xfer<T>(ctx, l - 1, v, fv.r, v.dmask);
                mg_smooth<T>(ctx, v, false, f, v.b, nullptr, nullptr, v.x, r, T(-1), 1, xf,
                             (double)sizeof(T) * (double)fv.nodes());
            } else if (s == 0) {
                mg_smooth<T>(ctx, v, l == 0 && upd, f, v.b, q, nullptr, v.x, r);
            } else {
                mg_smooth<T>(ctx, v, false, f, v.r, nullptr, v.x, v.x, r);
            }
            prof_close(ctx, h);
        }
        if (!last) {
            int h = open(cat(l, PF_KCAT_MG_RESIDUAL));
            mg_residual<T>(ctx, l, v);
            prof_close(ctx, h);
            if (!fuse) {
                const MgView<T> c = mg_view<T>(ctx, l + 1, r, z);
                const MgTransfer tr = mg_transfer(ctx, (size_t)l);
                h = open(cat(l, PF_KCAT_MG_TRANSFER));
                const int hk = prof_kern<T>(ctx, PF_K_RESTRICT, l,
                                            sizeof(T) * (double)(v.nodes() + c.nodes()) + mask_bytes(c.nodes()));
                pf_launch(ctx, (k_mg_restrict<T>), dim3(grid1(ctx->mg[l].N, 256, 1 << 20)), dim3(256), 0,
                                   ctx->stream, tr, c.rwx, c.rwy, (const T*)v.r, c.b, c.dmask, ctx->st);
                prof_close(ctx, hk);
                prof_close(ctx, h);
            }
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        if (l == 0) prof_close(ctx, hc);
        const MgView<T> v = mg_view<T>(ctx, l, r, z);
        const MgView<T> c = mg_view<T>(ctx, l + 1, r, z);
        if (fuse) {
            // r = b - A (x + P x_c), then x = (x + P x_c) + omega M^-1 r
            const MgXfer<T> xf = mg_xfer<T>(ctx, l, c, c.x, v.dmask);
            int h = open(cat(l, PF_KCAT_MG_RESIDUAL));
            mg_residual<T>(ctx, l, v, mg_prol(xf), c.nodes());
            prof_close(ctx, h);
            h = open(cat(l, PF_KCAT_MG_SMOOTH));
            mg_smooth<T>(ctx, v, false, l == 0 ? fin : 0, v.r, nullptr, v.x, v.x, r, T(-1), 2, xf,
                         (double)sizeof(T) * (double)c.nodes());
            prof_close(ctx, h);
            continue;
        }
        const MgTransfer tr = mg_transfer(ctx, (size_t)l);
        const long long nf = (long long)v.ncol * v.planes;
        int h = open(cat(l, PF_KCAT_MG_TRANSFER));
        const int hk = prof_kern<T>(ctx, PF_K_PROLONG, l, sizeof(T) * (double)(2 * nf + c.nodes()) + mask_bytes(nf));
        if (tr.n0f % 2 == 0)  // two fine nodes of a row per thread
            pf_launch(ctx, (k_mg_prolong_add2<T>), dim3(grid1(nf / 2, 256, 1 << 20)), dim3(256), 0, ctx->stream, tr,
                      c.wx, c.wy, (const T*)c.x, v.x, v.dmask, ctx->st);
        else
            pf_launch(ctx, (k_mg_prolong_add<T>), dim3(grid1(nf, 256, 1 << 20)), dim3(256), 0, ctx->stream, tr, c.wx,
                      c.wy, (const T*)c.x, v.x, v.dmask, ctx->st);
        prof_close(ctx, hk);
        prof_close(ctx, h);
        h = open(cat(l, PF_KCAT_MG_RESIDUAL));
        mg_residual<T>(ctx, l, v);
        prof_close(ctx, h);
        h = open(cat(l, PF_KCAT_MG_SMOOTH));
        mg_smooth<T>(ctx, v, false, l == 0 ? fin : 0, v.r, nullptr, v.x, v.x, r);
        prof_close(ctx, h);
    }
}

// One CG run on (x, r already holding b - A x).  T selects FP64 or FP32 vectors.
template <typename T>
static int run_cg(pf_ctx* ctx, int precond, T* x, T* r, T* p, T* p2, T* q, T* z, const T* dinv, const T* la,
                  const T* lw, const T* lcp, const T* kh, const T* kv, double stop2, int maxit, int check_every,
                  hipGraphExec_t* gcache, int* iters_out, int* status_out, double resume_rz = 0.) {
    const long long N = ctx->N;
    PcgState init{};
    init.stop2 = stop2;
    init.maxit = maxit;
    init.beta0 = resume_rz;  // > 0: keep the direction already in p (see finish_start)
    *ctx->hst = init;  // pinned staging: the host does not touch it again before the next sync
    HIPCK(hipMemcpyAsync(ctx->st, ctx->hst, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    const int ncol = (int)(ctx->dim == 3 ? ctx->n[0] * ctx->n[1] : ctx->n[0]);
    const int nplanes = (int)(ctx->dim == 3 ? ctx->n[2] : ctx->n[1]);
    const long long plane = ncol;
    const dim3 lgrid((ncol + 63) / 64, 1);
    const bool line = precond == 2, mg = precond == 3;
    // FP64 CG with the preconditioner in FP32: r -> r32, z32 = M^-1 r32, z = z32 + dots
    const bool p32 = sizeof(T) == 8 && ctx->prec32;
    auto precond32 = [&](bool start) {
        if constexpr (sizeof(T) == 8) {
            int hk = prof_kern<T>(ctx, PF_K_TO_F32, 0, (start ? 12.0 : 28.0) * N);
            pf_launch(ctx, k_cg_r_to_f32, dim3(grid1(N, 256, 4096)), dim3(256), 0, ctx->stream, N, !start,
                               (double*)r, (const double*)q, ctx->r32, ctx->st);
            prof_close(ctx, hk);
            if (mg) {
                mg_vcycle<float>(ctx, ctx->r32, nullptr, ctx->z32, false, 0);
            } else {  // exact line solve in FP32 (the smoother kernel with omega = 1)
                const MgView<float> v0 = mg_view<float>(ctx, 0, ctx->r32, ctx->z32);
                mg_smooth<float>(ctx, v0, false, 0, ctx->r32, nullptr, nullptr, ctx->z32, nullptr, 1.f);
            }
            hk = prof_kern<T>(ctx, PF_K_FROM_F32, 0, 20.0 * N);
            pf_launch(ctx, k_cg_z_from_f32, dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, start ? 2 : 1,
                               (const double*)r, (const float*)ctx->z32, (double*)z, ctx->st, ctx->partials);
            prof_close(ctx, hk);
        }
    };
    const int pm = sizeof(T) == 8 ? ctx->pm64 : ctx->pm32;
    const T* schur = (const T*)(sizeof(T) == 8 ? (const void*)ctx->schur : (const void*)ctx->schur32);
    PartShape ps = part_shape<T>(pm, nplanes, ncol);
    auto line_start = [&]() {
        if (pm)
            launch_part<T, true>(ctx, ps, ncol, nplanes, plane, r, nullptr, la, lw, lcp, schur, z);
        else
            hipLaunchKernelGGL((k_cg_line<T, true>), lgrid, dim3(64), 0, ctx->stream, ncol, nplanes, plane, nplanes,
                               r, (const T*)nullptr, la, lw, lcp, z, ctx->st, ctx->partials);
    };
    // a profiled solve also times the start preconditioner (with the first chunk)
    if (ctx->profile) ctx->prof_iter = 0;
    if (p32)
        precond32(true);
    else if (mg)
        mg_vcycle<T>(ctx, r, nullptr, z, false, 2);
    else if (line)
        line_start();
    else
        hipLaunchKernelGGL((k_cg_start<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, r, dinv, z, ctx->st,
                           ctx->partials);
    ctx->prof_iter = -1;
    const double tb = (double)sizeof(T);
    auto update = [&](T* pout) {
        (void)pout;
        int hk = -1;
        if (p32) {
            precond32(false);
        } else if (mg) {
            mg_vcycle<T>(ctx, r, q, z, true, 1);
        } else if (line && pm) {
            hk = prof_kern<T>(ctx, PF_K_LINE, 0, tb * (7.0 * N + 3.0 * std::max(ps.P - 1, 0) * (double)ncol));
            launch_part<T, false>(ctx, ps, ncol, nplanes, plane, r, q, la, lw, lcp, schur, z);
        } else if (line) {
            hk = prof_kern<T>(ctx, PF_K_LINE, 0, tb * 9.0 * N);
            pf_launch(ctx, (k_cg_line<T, false>), lgrid, dim3(64), 0, ctx->stream, ncol, nplanes, plane, nplanes,
                               r, (const T*)q, la, lw, lcp, z, ctx->st, ctx->partials);
        } else {
            hk = prof_kern<T>(ctx, PF_K_JACOBI, 0, tb * 5.0 * N);
            pf_launch(ctx, (k_cg_update<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, r, q, dinv, z,
                               ctx->st, ctx->partials);
        }
        prof_close(ctx, hk);
    };
    // a fresh run starts both direction buffers at zero (beta = 0 on the first iteration);
    // a resumed run keeps the previous direction in p
    if (!(resume_rz > 0.)) hipLaunchKernelGGL((k_fill<T>), dim3(grid1(N, 256, 4096)), dim3(256), 0, ctx->stream, N, p, T(0));
    hipLaunchKernelGGL((k_fill<T>), dim3(grid1(N, 256, 4096)), dim3(256), 0, ctx->stream, N, p2, T(0));
    HIPCK(hipGetLastError());
    // iteration t reads p_old from pb[t % 2] and writes p_new to pb[(t + 1) % 2]
    T* pb[2] = {p, p2};
    // p = z + beta p_old, x += alpha_prev p_old, q = A p, (p, q) and alpha: one fused apply,
    // or with a boundary matrix the same steps unfused (its rows are added between the
    // operator and the dot product)
    auto apply_step = [&](T* pin, T* pout) {
        if (!ctx->have_bnd) {
            launch_apply<T, MODE_OP, true, true>(ctx, nullptr, z, pin, pout, q, nullptr, kh, kv, x);
            return;
        }
        hipLaunchKernelGGL((k_cg_pstep<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, (const T*)z,
                           (const T*)pin, pout, x, (const PcgState*)ctx->st);
        launch_apply<T, MODE_OP, false, false>(ctx, pout, nullptr, nullptr, nullptr, q, nullptr, kh, kv);
        hipLaunchKernelGGL((k_cg_dot<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, (const T*)pout,
                           (const T*)q, (const uint64_t*)ctx->dmask, ctx->st, ctx->partials);
    };
    auto one_iter = [&](int t) {
        T* pin = pb[t & 1];
        T* pout = pb[(t + 1) & 1];
        apply_step(pin, pout);
        update(pout);
    };
    int chunk = check_every > 0 ? check_every : 32;
    if (mg) chunk = std::min(chunk, 4);  // few, expensive iterations: check the stop often
    chunk += chunk & 1;  // even: each graph replay starts on the same buffer parity
    if (ctx->profile) {
        // event-timed launches, no graph: events around every kernel of a chunk, one
        // synchronisation per chunk; only iterations that ran (not the early exits after
        // the stop) are counted
        int done_iters = 0;
        for (;;) {
            for (int c = 0; c < chunk; ++c) {
                T* pin = pb[c & 1];
                T* pout = pb[(c + 1) & 1];
                ctx->prof_iter = c;
                int h = prof_open(ctx, PF_KCAT_APPLY);
                const int hk = prof_kern<T>(ctx, PF_K_APPLY, 0, bytes_apply_fused(N, ctx->E, (int)sizeof(T)),
                                            !ctx->have_bnd);
                apply_step(pin, pout);
                prof_close(ctx, hk);
                prof_close(ctx, h);
                h = prof_open(ctx, PF_KCAT_PRECOND);
                update(pout);
                prof_close(ctx, h);
            }
            ctx->prof_iter = -1;
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
            HIPCK(hipStreamSynchronize(ctx->stream));
            prof_collect(ctx, std::min(chunk, ctx->hst->iters - done_iters));
            done_iters = ctx->hst->iters;
            if (ctx->hst->done) break;
        }
    } else if (!ctx->graphs) {
        // plain launches, one state read-back per chunk
        for (;;) {
            for (int c = 0; c < chunk; ++c) one_iter(c);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
            HIPCK(hipStreamSynchronize(ctx->stream));
            if (ctx->hst->done) break;
        }
    } else {
        // the cached graph of this precision holds one chunk of iterations for one
        // preconditioner set-up; anything else is captured anew
        const int gi = sizeof(T) == 8 ? 0 : 1;
        const int key = precond + (p32 ? 16 : 0);
        if (!*gcache || ctx->graph_chunk[gi] != chunk || ctx->graph_key[gi] != key) {
            if (*gcache) hipGraphExecDestroy(*gcache);
            *gcache = nullptr;
            hipGraph_t graph;
            HIPCK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            for (int c = 0; c < chunk; ++c) one_iter(c);
            HIPCK(hipStreamEndCapture(ctx->stream, &graph));
            HIPCK(hipGraphInstantiate(gcache, graph, nullptr, nullptr, 0));
            hipGraphDestroy(graph);
            ctx->graph_chunk[gi] = chunk;
            ctx->graph_key[gi] = key;
        }
        for (;;) {
            HIPCK(hipGraphLaunch(*gcache, ctx->stream));
            HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
            HIPCK(hipStreamSynchronize(ctx->stream));
            if (ctx->hst->done) break;
        }
    }
    // the last solution update (the fused apply of a further iteration would have done it)
    hipLaunchKernelGGL((k_x_fixup<T>), dim3(grid1(N, 256, 4096)), dim3(256), 0, ctx->stream, N, x, (const T*)p,
                       (const T*)p2, ctx->st);
    HIPCK(hipGetLastError());
    *iters_out = ctx->hst->iters;
    *status_out = ctx->hst->status;
    return PF_OK;
}

// dst = f * src (FP32; src may alias dst)
__global__ void k_scale_f32(long long n, const float* src, float f, float* dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = f * src[i];
}

// ||M(F - A0 x)||^2 + 0 (Dirichlet rows of x hold the values exactly) into r; returns sqrt
static int true_residual(pf_ctx* ctx, double* rn) {
    launch_apply<double, MODE_RES, false, false>(ctx, ctx->x, nullptr, nullptr, nullptr, ctx->r, ctx->F, ctx->kh, ctx->kv);
    hipLaunchKernelGGL((k_norm2<double>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, (long long)ctx->N, ctx->r,
                       ctx->st, ctx->partials, 1);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    *rn = std::sqrt(ctx->hst->rr);
    return PF_OK;
}

extern "C" {

int pf_solve(pf_ctx* ctx, const pf_solve_opts* opts, double* x_inout, pf_solve_stats* stats) {
    if (!ctx || !opts) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = prepare(ctx))) return rc;
    const long long N = ctx->N;
    const int gb = grid1(N, 256, 4096);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < 8; ++i) {
        ctx->kt[i] = 0.;
        ctx->kc[i] = 0;
    }
    ctx->kstat.clear();
    // start vector with the Dirichlet values
    if (opts->x0 == PF_X0_GIVEN) {
        if (!x_inout) return fail(ctx, PF_EINVAL, "x_inout required with PF_X0_GIVEN");
        HIPCK(hipMemcpyAsync(ctx->x, x_inout, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    } else if (opts->x0 == PF_X0_UNIFORM) {
        hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x, opts->init_value);
    } else if (opts->x0 == PF_X0_ZERO || !ctx->have_solution) {
        hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x, 0.);
    }
    if (ctx->nd)
        hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd, ctx->dnodes,
                           ctx->dvals, ctx->x);
    double bnorm = ctx->bnorm;
    if (opts->norm == PF_NORM_B_FULL) {
        bnorm = std::sqrt(ctx->bnorm * ctx->bnorm + ctx->dvals_sq);
    } else if (opts->norm == PF_NORM_INIT || opts->norm == PF_NORM_START) {
        hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->tmp, opts->init_value);
        if (ctx->nd)
            hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd,
                               ctx->dnodes, ctx->dvals, ctx->tmp);
        launch_apply<double, MODE_RES, false, false>(ctx, ctx->tmp, nullptr, nullptr, nullptr, ctx->qv, ctx->F, ctx->kh, ctx->kv);
        hipLaunchKernelGGL((k_norm2<double>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, ctx->qv, ctx->st,
                           ctx->partials, 1);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
        HIPCK(hipStreamSynchronize(ctx->stream));
        bnorm = std::sqrt(ctx->hst->rr);
    } else if (opts->norm != PF_NORM_B) {
        return fail(ctx, PF_EINVAL, "unknown residual norm");
    }
    if (!(bnorm > 0.)) bnorm = 1.;
    const double rtol = opts->rtol > 0. ? opts->rtol : 1e-8;
    const int precond = opts->precond == 1 ? 1 : opts->precond == 3 ? 3 : 2;  // 1 Jacobi, 2 (default) line, 3 multigrid
    if (precond == 3 && !ctx->mg_built && (rc = mg_build(ctx))) return rc;
    const int maxit = opts->maxit > 0 ? opts->maxit : 100000;
    int total_iters = 0, outer = 0, status = 0;
    double rn = 0.;
    if ((rc = true_residual(ctx, &rn))) return rc;
    if (opts->norm == PF_NORM_START) {
        if (opts->x0 == PF_X0_DEVICE_CHECKED) return fail(ctx, PF_EINVAL, "PF_X0_DEVICE_CHECKED needs PF_NORM_INIT");
        // the floor keeps the target rtol * bnorm at or above 1e-13 ||B - A X_init||, which
        // FP64 CG attains (a converged Picard loop starts from a residual near zero)
        bnorm = std::max(rn, std::max(1e-4, 1e-13 / std::max(rtol, 1e-300)) * bnorm);
        if (!(bnorm > 0.)) bnorm = 1.;
    }
    int start = 0;
    if (opts->x0 == PF_X0_DEVICE_CHECKED) {
        if (opts->norm != PF_NORM_INIT) return fail(ctx, PF_EINVAL, "PF_X0_DEVICE_CHECKED needs PF_NORM_INIT");
        if (!(rn <= bnorm)) {  // the given start is no better than the uniform one (or not finite)
            hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x, opts->init_value);
            if (ctx->nd)
                hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd,
                                   ctx->dnodes, ctx->dvals, ctx->x);
            if ((rc = true_residual(ctx, &rn))) return rc;
            start = 1;
        }
    }
    // FP64 CG from the current x (r = the true residual); restarts only if the recursive
    // residual drifted from the true one
    auto fp64_cg = [&](bool prec32) -> int {
        for (int attempt = 0; attempt < 4 && rn > rtol * bnorm && total_iters < maxit; ++attempt) {
            int it = 0, stt = 0;
            double stop2 = (rtol * bnorm) * (rtol * bnorm);
            ctx->prec32 = prec32;
            int rc2 = run_cg<double>(ctx, precond, ctx->x, ctx->r, ctx->p, ctx->p2, ctx->qv, ctx->z, ctx->dinv, ctx->la,
                                     ctx->lw, ctx->lcp, ctx->kh, ctx->kv, stop2, maxit - total_iters, opts->check_every,
                                     &ctx->graph64, &it, &stt);
            ctx->prec32 = false;
            if (rc2) return rc2;
            total_iters += it;
            status = stt;
            if ((rc2 = true_residual(ctx, &rn))) return rc2;
            if (stt != 0) break;
            outer++;
        }
        return PF_OK;
    };
    // FP32 forms: iterative refinement converges while the FP32 correction contracts the
    // FP64 residual (roughly kappa(A) * eps_fp32 < 1); a step that fails to halve the true
    // residual means it no longer does, and the solve finishes with FP64 CG
    int fallback = 0;
    auto stalled = [&](double before) { return rn > rtol * bnorm && rn > 0.5 * before; };
    if (opts->precision == PF_FP64 || opts->precision == PF_FP64_PREC32) {
        if (opts->precision == PF_FP64_PREC32 && (precond == 1 || !ctx->pm32))
            return fail(ctx, PF_EINVAL, "PF_FP64_PREC32 needs the partitioned line or the multigrid preconditioner");
        if ((rc = fp64_cg(opts->precision == PF_FP64_PREC32))) return rc;
    } else if (opts->precision == PF_FP32_REFINE) {
        const double inner = opts->inner_rtol > 0. ? opts->inner_rtol : 1e-4;
        const int max_outer = opts->max_outer > 0 ? opts->max_outer : 50;
        while (rn > rtol * bnorm && outer < max_outer && total_iters < maxit) {
            const double before = rn;
            // r holds M(F - A0 x) in FP64; scale to unit norm and solve A d = r in FP32
            double s = 1. / rn;
            hipLaunchKernelGGL(k_scale_to_f32, dim3(gb), dim3(256), 0, ctx->stream, N, ctx->r, s, ctx->r32);
            hipLaunchKernelGGL((k_fill<float>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x32, 0.f);
            int it = 0, stt = 0;
            // the target for this step: reduce by `inner`, but not beyond what the outer tolerance needs
            double goal = std::max(inner, 0.5 * rtol * bnorm / rn);
            if ((rc = run_cg<float>(ctx, precond, ctx->x32, ctx->r32, ctx->p32, ctx->p2_32, ctx->q32, ctx->z32, ctx->dinv32,
                                    ctx->la32, ctx->lw32, ctx->lcp32, ctx->kh32, ctx->kv32, goal * goal, maxit - total_iters, opts->check_every, &ctx->graph32, &it,
                                    &stt)))
                return rc;
            total_iters += it;
            hipLaunchKernelGGL((k_add_correction<float>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x, ctx->x32, rn);
            HIPCK(hipGetLastError());
            outer++;
            if ((rc = true_residual(ctx, &rn))) return rc;
            if (stt < 0) {
                status = stt;
                break;
            }
            if (stalled(before)) {
                fallback = 1;
                break;
            }
        }
    } else if (opts->precision == PF_FP32_RELIABLE) {
        // One FP32 Krylov run with residual replacement (reliable updates): whenever the
        // FP32 recursive residual has dropped by `inner`, the correction is added to x in
        // FP64, the residual is recomputed in FP64 and CG continues with the previous
        // direction, rescaled to the new residual's units, instead of restarting.  The
        // direction is kept only while the true residual stays close to the recursive one
        // (within a factor 3): the FP32 operator's error grows with the reduction asked of
        // one run, so `inner` should be modest (1e-2 to 1e-3) for this form.
        const double inner = opts->inner_rtol > 0. ? opts->inner_rtol : 1e-2;
        const int max_outer = opts->max_outer > 0 ? opts->max_outer : 50;
        double rn_prev = 0., rz_last = 0.;
        int last_it = 0;
        bool keep = false;  // resume with the kept direction (else a fresh start)
        while (rn > rtol * bnorm && outer < max_outer && total_iters < maxit) {
            double s = 1. / rn;
            hipLaunchKernelGGL(k_scale_to_f32, dim3(gb), dim3(256), 0, ctx->stream, N, ctx->r, s, ctx->r32);
            hipLaunchKernelGGL((k_fill<float>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x32, 0.f);
            double resume = 0.;
            if (keep && rz_last > 0.) {
                const double f = rn_prev / rn;  // old scale over new scale
                const float* src = (last_it & 1) ? ctx->p2_32 : ctx->p32;
                hipLaunchKernelGGL(k_scale_f32, dim3(gb), dim3(256), 0, ctx->stream, N, src, (float)f, ctx->p32);
                resume = rz_last * f * f;
            }
            int it = 0, stt = 0;
            double goal = std::max(inner, 0.5 * rtol * bnorm / rn);
            if ((rc = run_cg<float>(ctx, precond, ctx->x32, ctx->r32, ctx->p32, ctx->p2_32, ctx->q32, ctx->z32,
                                    ctx->dinv32, ctx->la32, ctx->lw32, ctx->lcp32, ctx->kh32, ctx->kv32, goal * goal,
                                    maxit - total_iters, opts->check_every, &ctx->graph32, &it, &stt, resume)))
                return rc;
            total_iters += it;
            last_it = it;
            rz_last = ctx->hst->rz_prev;  // (r, z) the last direction was built from
            const double rec = std::sqrt(ctx->hst->rr);  // recursive residual, units of rn
            rn_prev = rn;
            hipLaunchKernelGGL((k_add_correction<float>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->x, ctx->x32, rn);
            HIPCK(hipGetLastError());
            outer++;
            if ((rc = true_residual(ctx, &rn))) return rc;
            if (stt < 0) {
                status = stt;
                break;
            }
            if (stalled(rn_prev)) {
                fallback = 1;
                break;
            }
            // a direction is kept only if the replaced residual is close to the recursive
            // one; otherwise it is no longer conjugate to the new residual's error and the
            // next run restarts from the preconditioned residual
            keep = it > 0 && rn < 0.5 * rn_prev && rn <= 3. * rec * rn_prev;
        }
    } else {
        return fail(ctx, PF_EINVAL, "unknown precision");
    }
    if (fallback && status == 0 && (rc = fp64_cg(false))) return rc;
    HIPCK(hipStreamSynchronize(ctx->stream));
    auto t1 = std::chrono::steady_clock::now();
    ctx->have_solution = true;
    if (x_inout) HIPCK(hipMemcpy(x_inout, ctx->x, sizeof(double) * N, hipMemcpyDeviceToHost));
    int ret = PF_OK;
    if (status == -6 || status == -7) ret = PF_EBREAK;
    else if (status == -8) ret = PF_ENONFINITE;
    else if (rn > rtol * bnorm) ret = PF_ENOCONV;
    if (stats) {
        stats->iters = total_iters;
        stats->outer = outer;
        stats->rel_resid = rn / bnorm;
        stats->bnorm = bnorm;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->status = ret;
        stats->start = start;
        stats->fallback = fallback;
    }
    if (ret == PF_EBREAK) ctx->err = "CG breakdown: stiffness matrix is not positive definite";
    if (ret == PF_ENONFINITE) ctx->err = "non-finite value in CG iteration";
    if (ret == PF_ENOCONV) ctx->err = "did not converge to the requested tolerance";
    return ret;
}

int pf_get_solution(pf_ctx* ctx, double* x_out) {
    if (!ctx || !x_out) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no solution on the device");
    HIPCK(hipSetDevice(ctx->device));
    HIPCK(hipMemcpy(x_out, ctx->x, sizeof(double) * ctx->N, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_apply(pf_ctx* ctx, int precision, const void* x_dev, void* y_dev) {
    if (!ctx || !x_dev || !y_dev) return PF_EINVAL;
    if (!ctx->have_coeffs) return fail(ctx, PF_ESTATE, "pf_coeffs_set first");
    HIPCK(hipSetDevice(ctx->device));
    if (precision == PF_FP64)
        launch_apply<double, MODE_OP, false, false>(ctx, (const double*)x_dev, nullptr, nullptr, nullptr, (double*)y_dev, nullptr,
                                                    ctx->kh, ctx->kv);
    else
        launch_apply<float, MODE_OP, false, false>(ctx, (const float*)x_dev, nullptr, nullptr, nullptr, (float*)y_dev, nullptr,
                                                   ctx->kh32, ctx->kv32);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_apply_host(pf_ctx* ctx, const double* x, double* y) {
    if (!ctx || !x || !y) return PF_EINVAL;
    if (!ctx->have_coeffs) return fail(ctx, PF_ESTATE, "pf_coeffs_set first");
    HIPCK(hipSetDevice(ctx->device));
    HIPCK(hipMemcpy(ctx->tmp, x, sizeof(double) * ctx->N, hipMemcpyHostToDevice));
    launch_apply<double, MODE_OP, false, false>(ctx, ctx->tmp, nullptr, nullptr, nullptr, ctx->qv, nullptr, ctx->kh, ctx->kv);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipMemcpy(y, ctx->qv, sizeof(double) * ctx->N, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_precond_apply_host(pf_ctx* ctx, int precision, int precond, const double* r_in, double* z_out) {
    if (!ctx || !r_in || !z_out) return PF_EINVAL;
    if (precision != PF_FP64 && precision != PF_FP32_REFINE && precision != PF_FP32_RELIABLE)
        return fail(ctx, PF_EINVAL, "unknown precision");
    if (precond != 1 && precond != 2 && precond != 3) return fail(ctx, PF_EINVAL, "precond must be 1, 2 or 3");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = prepare(ctx))) return rc;
    if (precond == 3 && !ctx->mg_built && (rc = mg_build(ctx))) return rc;
    const long long N = ctx->N;
    const int gb = grid1(N, 256, 4096);
    PcgState init{};
    init.maxit = 1;
    *ctx->hst = init;
    HIPCK(hipMemcpyAsync(ctx->st, ctx->hst, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->tmp, r_in, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    const int ncol = (int)(ctx->dim == 3 ? ctx->n[0] * ctx->n[1] : ctx->n[0]);
    const int nplanes = (int)(ctx->dim == 3 ? ctx->n[2] : ctx->n[1]);
    auto run = [&](auto* r, auto* z, auto* dinv, auto* la, auto* lw, auto* lcp, auto* schur, int pm) {
        using T = std::remove_pointer_t<decltype(r)>;
        if (precond == 3) {
            mg_vcycle<T>(ctx, r, nullptr, z, false, 2);
        } else if (precond == 2 && pm) {
            launch_part<T, true>(ctx, part_shape<T>(pm, nplanes, ncol), ncol, nplanes, ncol, r, nullptr, la, lw, lcp,
                                 schur, z);
        } else if (precond == 2) {
            hipLaunchKernelGGL((k_cg_line<T, true>), dim3((ncol + 63) / 64), dim3(64), 0, ctx->stream, ncol, nplanes,
                               (long long)ncol, nplanes, r, (const T*)nullptr, la, lw, lcp, z, ctx->st, ctx->partials);
        } else {
            hipLaunchKernelGGL((k_cg_start<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, r, dinv, z,
                               ctx->st, ctx->partials);
        }
    };
    if (precision == PF_FP64) {
        HIPCK(hipMemcpyAsync(ctx->r, ctx->tmp, sizeof(double) * N, hipMemcpyDeviceToDevice, ctx->stream));
        run(ctx->r, ctx->z, ctx->dinv, ctx->la, ctx->lw, ctx->lcp, ctx->schur, ctx->pm64);
        HIPCK(hipMemcpyAsync(ctx->tmp, ctx->z, sizeof(double) * N, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        hipLaunchKernelGGL((k_convert<float, double>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->tmp, ctx->r32);
        run(ctx->r32, ctx->z32, ctx->dinv32, ctx->la32, ctx->lw32, ctx->lcp32, ctx->schur32, ctx->pm32);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gb), dim3(256), 0, ctx->stream, N, ctx->z32, ctx->tmp);
    }
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(z_out, ctx->tmp, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_rhs_host(pf_ctx* ctx, double* b_out) {
    if (!ctx || !b_out) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    ctx->dirty = true;
    if ((rc = prepare(ctx))) return rc;  // leaves M(F - A0 xD) in r
    std::vector<double> h(ctx->N);
    HIPCK(hipMemcpy(b_out, ctx->r, sizeof(double) * ctx->N, hipMemcpyDeviceToHost));
    if (ctx->nd) {
        std::vector<long long> nodes(ctx->nd);
        std::vector<double> vals(ctx->nd);
        HIPCK(hipMemcpy(nodes.data(), ctx->dnodes, sizeof(long long) * ctx->nd, hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(vals.data(), ctx->dvals, sizeof(double) * ctx->nd, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < ctx->nd; ++i) b_out[nodes[i]] = vals[i];
    }
    return PF_OK;
}

int pf_kernel_stats(const pf_ctx* ctx, pf_kernel_stat* out, int nmax) {
    if (!ctx || nmax < 0 || (nmax > 0 && !out)) return PF_EINVAL;
    const int n = (int)std::min<size_t>((size_t)nmax, ctx->kstat.size());
    for (int i = 0; i < n; ++i) out[i] = ctx->kstat[i];
    return (int)ctx->kstat.size();
}

int pf_kernel_times(pf_ctx* ctx, double* seconds, int64_t* counts, int n) {
    if (!ctx) return PF_EINVAL;
    for (int i = 0; i < n && i < 8; ++i) {
        if (seconds) seconds[i] = ctx->kc[i] ? ctx->kt[i] / ctx->kc[i] : 0.;
        if (counts) counts[i] = ctx->kc[i];
    }
    return PF_OK;
}

static ElemGrid egrid(pf_ctx* ctx) {
    return ElemGrid{ctx->dim, (int)ctx->n[0], (int)ctx->n[1], (int)ctx->n[2]};
}

}  // extern "C"

// Trilinear interpolation of node values of a coarser rectilinear grid onto the
// solution vector: per fine index along each axis, the coarse interval (lo) and
// the weight of its upper end.  RAW: the coarse values are raw network outputs c and
// the node values offset + scale c are formed here, rounded to float as the host does
// for the RAW = false form (separate multiply and add, no fused contraction), so both
// forms give identical results.
template <bool RAW>
__global__ void k_prolong3d(int nx, int ny, int nz, int cx, int cy, const float* __restrict__ src, double offset,
                            double scale, const int* __restrict__ ix, const int* __restrict__ iy,
                            const int* __restrict__ iz, const float* __restrict__ wx, const float* __restrict__ wy,
                            const float* __restrict__ wz, double* __restrict__ x) {
    const long long n = (long long)nx * ny * nz;
    auto at = [&](long long i) -> float {
        if (RAW) return (float)__dadd_rn(offset, __dmul_rn(scale, (double)src[i]));
        return src[i];
    };
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long long)nx * ny));
        const int a = ix[i], b = iy[j], c = iz[k];
        const float u = wx[i], w = wy[j], t = wz[k];
        const long long base = ((long long)c * cy + b) * cx + a;
        const long long sy = cx, sz = (long long)cx * cy;
        const float s000 = at(base), s100 = at(base + 1), s010 = at(base + sy), s110 = at(base + sy + 1);
        const float s001 = at(base + sz), s101 = at(base + sz + 1), s011 = at(base + sz + sy),
                    s111 = at(base + sz + sy + 1);
        const float v00 = s000 + u * (s100 - s000);
        const float v10 = s010 + u * (s110 - s010);
        const float v01 = s001 + u * (s101 - s001);
        const float v11 = s011 + u * (s111 - s011);
        const float lo = v00 + w * (v10 - v00), hi = v01 + w * (v11 - v01);
        x[v] = (double)(lo + t * (hi - lo));
    }
}

// Per fine node of each axis: the coarse interval and the weight of its upper end
// (trilinear in physical coordinates; fine nodes outside the coarse axis clamp).
static int prolong_tables(pf_ctx* ctx, const double* cx, int ncx, const double* cy, int ncy, const double* cz,
                          int ncz, std::vector<int> idx[3], std::vector<float> wt[3]) {
    if (ctx->dim != 3) return fail(ctx, PF_ESHAPE, "solution prolongation: 3D grids only");
    if (ncx < 2 || ncy < 2 || ncz < 2) return fail(ctx, PF_ESHAPE, "coarse grid needs >= 2 nodes per axis");
    const double* ca[3] = {cx, cy, cz};
    const int nc[3] = {ncx, ncy, ncz};
    for (int a = 0; a < 3; ++a) {
        for (int i = 0; i + 1 < nc[a]; ++i)
            if (!(ca[a][i + 1] > ca[a][i])) return fail(ctx, PF_EINVAL, "coarse axis coordinates must increase");
        const std::vector<double>& f = ctx->ax[a];
        idx[a].resize(f.size());
        wt[a].resize(f.size());
        for (size_t i = 0; i < f.size(); ++i) {
            const double* c = ca[a];
            int lo = (int)(std::upper_bound(c, c + nc[a], f[i]) - c) - 1;
            lo = std::max(0, std::min(lo, nc[a] - 2));
            double w = (f[i] - c[lo]) / (c[lo + 1] - c[lo]);
            idx[a][i] = lo;
            wt[a][i] = (float)std::max(0., std::min(1., w));
        }
    }
    return PF_OK;
}

extern "C" {

int pf_solution_prolong(pf_ctx* ctx, const float* coarse, const double* cx, int ncx, const double* cy, int ncy,
                        const double* cz, int ncz, double offset, double scale) {
    if (!ctx || !coarse || !cx || !cy || !cz) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    std::vector<int> idx[3];
    std::vector<float> wt[3];
    int rc;
    if ((rc = prolong_tables(ctx, cx, ncx, cy, ncy, cz, ncz, idx, wt))) return rc;
    const size_t ncoarse = (size_t)ncx * ncy * ncz;
    std::vector<float> src(ncoarse);
    for (size_t i = 0; i < ncoarse; ++i) src[i] = (float)(offset + scale * (double)coarse[i]);
    const size_t nidx = ctx->ax[0].size() + ctx->ax[1].size() + ctx->ax[2].size();
    float* dsrc = nullptr;
    char* dtab = nullptr;
    HIPCK(hipMalloc((void**)&dsrc, sizeof(float) * ncoarse));
    if (hipMalloc((void**)&dtab, (sizeof(int) + sizeof(float)) * nidx) != hipSuccess) {
        hipFree(dsrc);
        return fail(ctx, PF_EDEVICE, "pf_solution_prolong: out of device memory");
    }
    int* di = (int*)dtab;
    float* dw = (float*)(dtab + sizeof(int) * nidx);
    size_t o = 0;
    int* dix[3];
    float* dwx[3];
    for (int a = 0; a < 3; ++a) {
        dix[a] = di + o;
        dwx[a] = dw + o;
        hipMemcpyAsync(dix[a], idx[a].data(), sizeof(int) * idx[a].size(), hipMemcpyHostToDevice, ctx->stream);
        hipMemcpyAsync(dwx[a], wt[a].data(), sizeof(float) * wt[a].size(), hipMemcpyHostToDevice, ctx->stream);
        o += idx[a].size();
    }
    hipMemcpyAsync(dsrc, src.data(), sizeof(float) * ncoarse, hipMemcpyHostToDevice, ctx->stream);
    hipLaunchKernelGGL((k_prolong3d<false>), dim3(grid1(ctx->N, 256, 8192)), dim3(256), 0, ctx->stream,
                       (int)ctx->n[0], (int)ctx->n[1], (int)ctx->n[2], ncx, ncy, dsrc, 0., 1., dix[0], dix[1], dix[2],
                       dwx[0], dwx[1], dwx[2], ctx->x);
    if (ctx->nd)
        hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd, ctx->dnodes,
                           ctx->dvals, ctx->x);
    hipError_t err = hipStreamSynchronize(ctx->stream);
    hipFree(dsrc);
    hipFree(dtab);
    if (err != hipSuccess) return fail(ctx, PF_EDEVICE, std::string("pf_solution_prolong: ") + hipGetErrorString(err));
    ctx->have_solution = true;
    return PF_OK;
}

int pf_solution_prolong_dev(pf_ctx* ctx, const float* coarse_dev, const double* cx, int ncx, const double* cy, int ncy,
                            const double* cz, int ncz, double offset, double scale) {
    if (!ctx || !coarse_dev || !cx || !cy || !cz) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    void* ws = scratch_ptr[ctx->device];
    if (!ws) return fail(ctx, PF_ESTATE, "pf_solution_prolong_dev: no workspace address for this device");
    const size_t nidx = ctx->ax[0].size() + ctx->ax[1].size() + ctx->ax[2].size();
    if ((sizeof(int) + sizeof(float)) * nidx > PF_PROLONG_SCRATCH_BYTES)
        return fail(ctx, PF_ESHAPE, "pf_solution_prolong_dev: the grid's axes have " + std::to_string(nidx) +
                                        " nodes in all; at most " +
                                        std::to_string(PF_PROLONG_SCRATCH_BYTES / (sizeof(int) + sizeof(float))) +
                                        " fit the fixed workspace");
    std::vector<int> idx[3];
    std::vector<float> wt[3];
    int rc;
    if ((rc = prolong_tables(ctx, cx, ncx, cy, ncy, cz, ncz, idx, wt))) return rc;
    int* di = (int*)ws;
    float* dw = (float*)((char*)ws + sizeof(int) * nidx);
    size_t o = 0;
    int* dix[3];
    float* dwx[3];
    for (int a = 0; a < 3; ++a) {
        dix[a] = di + o;
        dwx[a] = dw + o;
        HIPCK(hipMemcpyAsync(dix[a], idx[a].data(), sizeof(int) * idx[a].size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCK(hipMemcpyAsync(dwx[a], wt[a].data(), sizeof(float) * wt[a].size(), hipMemcpyHostToDevice, ctx->stream));
        o += idx[a].size();
    }
    hipLaunchKernelGGL((k_prolong3d<true>), dim3(grid1(ctx->N, 256, 8192)), dim3(256), 0, ctx->stream,
                       (int)ctx->n[0], (int)ctx->n[1], (int)ctx->n[2], ncx, ncy, coarse_dev, offset, scale, dix[0],
                       dix[1], dix[2], dwx[0], dwx[1], dwx[2], ctx->x);
    if (ctx->nd)
        hipLaunchKernelGGL(k_scatter, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd, ctx->dnodes,
                           ctx->dvals, ctx->x);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));  // the host tables must outlive the copies
    ctx->have_solution = true;
    return PF_OK;
}

int pf_solution_set(pf_ctx* ctx, const double* x) {
    if (!ctx || !x) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    HIPCK(hipSetDevice(ctx->device));
    HIPCK(hipMemcpy(ctx->x, x, sizeof(double) * ctx->N, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->xprev, ctx->x, sizeof(double) * ctx->N, hipMemcpyDeviceToDevice));
    ctx->have_solution = true;
    return PF_OK;
}

int pf_solution_change(pf_ctx* ctx, double* max_change, double* max_value) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no solution on the device");
    HIPCK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_delta_stats, dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, (long long)ctx->N, ctx->x,
                       ctx->xprev, ctx->dout, ctx->partials, &ctx->st->counter[3]);
    HIPCK(hipGetLastError());
    double h[2];
    HIPCK(hipMemcpyAsync(h, ctx->dout, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (max_change) *max_change = h[0];
    if (max_value) *max_value = h[1];
    return PF_OK;
}

int pf_element_means(pf_ctx* ctx, double* out) {
    if (!ctx || !out) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no solution on the device");
    HIPCK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_elem_mean, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, egrid(ctx),
                       (long long)ctx->E, ctx->x, ctx->emean);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, ctx->emean, sizeof(double) * ctx->E, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_coeffs_powerlaw(pf_ctx* ctx, int equation, const double* k0_h, const double* k0_v, const double* a_h,
                       const double* a_v, int64_t nelem) {
    if (!ctx || !k0_h || !k0_v || !a_h || !a_v) return PF_EINVAL;
    if (nelem != ctx->E) return fail(ctx, PF_EINVAL, "law arrays must have one value per element");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if (!ctx->k0h) {
        if ((rc = dalloc(ctx, &ctx->k0h, ctx->E))) return rc;
        if ((rc = dalloc(ctx, &ctx->k0v, ctx->E))) return rc;
        if ((rc = dalloc(ctx, &ctx->ah, ctx->E))) return rc;
        if ((rc = dalloc(ctx, &ctx->av, ctx->E))) return rc;
    }
    HIPCK(hipMemcpy(ctx->k0h, k0_h, sizeof(double) * nelem, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->k0v, k0_v, sizeof(double) * nelem, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->ah, a_h, sizeof(double) * nelem, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->av, a_v, sizeof(double) * nelem, hipMemcpyHostToDevice));
    ctx->equation = equation;
    ctx->have_law = true;
    return PF_OK;
}

// the law of ctx at the element means of the nodal field t (ctx's grid)
static int coeffs_from(pf_ctx* ctx, const double* t) {
    HIPCK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_powerlaw, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, egrid(ctx),
                       (long long)ctx->E, t, ctx->k0h, ctx->k0v, ctx->ah, ctx->av, ctx->kh, ctx->kv, ctx->kh32,
                       ctx->kv32);
    dyn_capture(ctx);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->have_coeffs = true;
    ctx->dirty = true;
    return PF_OK;
}

int pf_coeffs_update(pf_ctx* ctx) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->have_law) return fail(ctx, PF_ESTATE, "pf_coeffs_powerlaw first");
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no solution (temperature) on the device");
    return coeffs_from(ctx, ctx->x);
}

// two contexts hold the same grid on the same device (element fields and nodal fields of
// one are laid out as the other's)
static bool same_grid(const pf_ctx* a, const pf_ctx* b) {
    if (a->device != b->device || a->dim != b->dim || a->N != b->N || a->E != b->E || a->geom != b->geom) return false;
    for (int d = 0; d < a->dim; ++d)
        if (a->ax[d] != b->ax[d]) return false;
    return true;
}

int pf_coeffs_update_from(pf_ctx* ctx, const pf_ctx* src) {
    if (!ctx || !src) return PF_EINVAL;
    if (!ctx->have_law) return fail(ctx, PF_ESTATE, "pf_coeffs_powerlaw first");
    if (!same_grid(ctx, src)) return fail(ctx, PF_EINVAL, "the temperature context holds another grid or device");
    if (!src->have_solution) return fail(ctx, PF_ESTATE, "no solution (temperature) in the temperature context");
    HIPCK(hipSetDevice(src->device));
    HIPCK(hipStreamSynchronize(src->stream));
    return coeffs_from(ctx, src->x);
}

int pf_source_from(pf_ctx* dst, const pf_ctx* src, int what) {
    if (!dst || !src) return PF_EINVAL;
    pf_ctx* ctx = dst;  // error reports go to the destination context
    if (what != PF_POST_JOULE) return fail(dst, PF_EINVAL, "pf_source_from: only PF_POST_JOULE");
    if (!same_grid(dst, src)) return fail(dst, PF_EINVAL, "the source context holds another grid or device");
    if (!src->have_cur) return fail(dst, PF_ESTATE, "the source context has no Joule heat (pf_currents first)");
    HIPCK(hipSetDevice(dst->device));
    HIPCK(hipStreamSynchronize(src->stream));
    HIPCK(hipMemcpyAsync(dst->q, src->heat, sizeof(double) * dst->E, hipMemcpyDeviceToDevice, dst->stream));
    HIPCK(hipStreamSynchronize(dst->stream));
    dst->have_source = true;
    dst->dirty = true;
    return PF_OK;
}

int pf_get_coeffs(pf_ctx* ctx, double* k_inplane, double* k_vert) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    // the conductivities as set (with a transient operator, kh / kv hold theta times them)
    const double* kh = ctx->dyn ? ctx->khB : ctx->kh;
    const double* kv = ctx->dyn ? ctx->kvB : ctx->kv;
    if (k_inplane) HIPCK(hipMemcpy(k_inplane, kh, sizeof(double) * ctx->E, hipMemcpyDeviceToHost));
    if (k_vert) HIPCK(hipMemcpy(k_vert, kv, sizeof(double) * ctx->E, hipMemcpyDeviceToHost));
    return PF_OK;
}

// dsh = capacity (transient) + boundary-matrix diagonal; off when neither is set
static int rebuild_dsh(pf_ctx* ctx) {
    ctx->dsh_on = ctx->dyn || ctx->have_dconv;
    if (!ctx->dsh_on) return PF_OK;
    int rc;
    if (!ctx->dsh && ((rc = dalloc(ctx, &ctx->dsh, ctx->N)) || (rc = dalloc(ctx, &ctx->dsh32, ctx->N)))) return rc;
    hipLaunchKernelGGL(k_sum_dsh, dim3(grid1(ctx->N, 256, 8192)), dim3(256), 0, ctx->stream, (long long)ctx->N,
                       (const double*)(ctx->dyn ? ctx->dcap : nullptr),
                       (const double*)(ctx->have_dconv ? ctx->dconv : nullptr), ctx->dsh, ctx->dsh32);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_boundary_matrix_set(pf_ctx* ctx, const int64_t* rows, const int64_t* cols, const double* vals, int64_t nnz) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (nnz < 0 || (nnz > 0 && (!rows || !cols || !vals))) return fail(ctx, PF_EINVAL, "bad boundary matrix");
    if (nnz > 0 && ctx->dyn) return fail(ctx, PF_ESTATE, "a boundary matrix with a transient operator is not supported");
    const int64_t N = ctx->N;
    std::vector<double> diag;
    std::map<std::pair<int64_t, int64_t>, double> off;  // (row, col), both triangles
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t r = rows[k], c = cols[k];
        if (r < 0 || r >= N || c < 0 || c >= N) return fail(ctx, PF_EINVAL, "boundary matrix index out of range");
        if (c > r) return fail(ctx, PF_EINVAL, "boundary matrix entries must be in the lower triangle (col <= row)");
        if (!std::isfinite(vals[k])) return fail(ctx, PF_EINVAL, "boundary matrix value not finite");
        if (r == c) {
            if (diag.empty()) diag.assign(N, 0.);
            diag[r] += vals[k];
        } else {
            off[{r, c}] += vals[k];
            off[{c, r}] += vals[k];
        }
    }
    HIPCK(hipSetDevice(ctx->device));
    HIPCK(hipStreamSynchronize(ctx->stream));
    int rc;
    ctx->have_dconv = !diag.empty();
    if (ctx->have_dconv) {
        if ((rc = dalloc(ctx, &ctx->dconv, N))) return rc;
        HIPCK(hipMemcpy(ctx->dconv, diag.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    }
    std::vector<long long> rr, cc;
    std::vector<int> ptr{0};
    std::vector<double> vv;
    for (const auto& kv : off) {
        if (rr.empty() || rr.back() != kv.first.first) {
            if (!rr.empty()) ptr.push_back((int)cc.size());
            rr.push_back(kv.first.first);
        }
        cc.push_back(kv.first.second);
        vv.push_back(kv.second);
    }
    if (!rr.empty()) ptr.push_back((int)cc.size());
    ctx->nbr = (long long)rr.size();
    ctx->have_bnd = ctx->nbr > 0;
    if (ctx->have_bnd) {
        std::vector<float> v32(vv.begin(), vv.end());
        if ((rc = upload(ctx, &ctx->bnd_rows, rr)) || (rc = upload(ctx, &ctx->bnd_ptr, ptr)) ||
            (rc = upload(ctx, &ctx->bnd_cols, cc)) || (rc = upload(ctx, &ctx->bnd_vals, vv)) ||
            (rc = upload(ctx, &ctx->bnd_vals32, v32)))
            return rc;
    }
    if ((rc = rebuild_dsh(ctx))) return rc;
    ctx->dirty = true;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_capacity_set(pf_ctx* ctx, const double* c_elem, int64_t nelem, double theta) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    HIPCK(hipSetDevice(ctx->device));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (!c_elem) {  // back to the steady operator
        if (ctx->dyn && ctx->have_coeffs) {
            hipLaunchKernelGGL(k_theta_scale, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream,
                               (long long)ctx->E, 1., (const double*)ctx->khB, (const double*)ctx->kvB, ctx->kh,
                               ctx->kv, ctx->kh32, ctx->kv32);
            HIPCK(hipStreamSynchronize(ctx->stream));
        }
        ctx->dyn = false;
        int rc = rebuild_dsh(ctx);
        ctx->dirty = true;
        drop_graphs(ctx);
        return rc;
    }
    if (ctx->have_dconv || ctx->have_bnd)
        return fail(ctx, PF_ESTATE, "a transient operator with a boundary matrix is not supported (the reference's "
                                    "transient solvers take temperature conditions only)");
    if (nelem != ctx->E) return fail(ctx, PF_EINVAL, "capacity array must have one value per element");
    if (!(theta >= 0. && theta <= 1.)) return fail(ctx, PF_EINVAL, "theta (methodparam) must be in [0, 1]");
    for (int64_t e = 0; e < nelem; ++e)
        if (!std::isfinite(c_elem[e]) || c_elem[e] < 0.)
            return fail(ctx, PF_EINVAL, "capacities must be finite and >= 0 (element " + std::to_string(e) + ")");
    int rc;
    const size_t N = ctx->N;
    if (!ctx->dcap) {
        if ((rc = dalloc(ctx, &ctx->dcap, N)) || (rc = dalloc(ctx, &ctx->F0, N)) ||
            (rc = dalloc(ctx, &ctx->khB, ctx->E)) || (rc = dalloc(ctx, &ctx->kvB, ctx->E)) ||
            (rc = dalloc(ctx, &ctx->zmask, (N + 63) / 64)))
            return rc;
        HIPCK(hipMemset(ctx->zmask, 0, sizeof(uint64_t) * ((N + 63) / 64)));
    }
    double* ce = nullptr;
    HIPCK(hipMalloc((void**)&ce, sizeof(double) * ctx->E));
    hipError_t e1 = hipMemcpyAsync(ce, c_elem, sizeof(double) * ctx->E, hipMemcpyHostToDevice, ctx->stream);
    if (e1 == hipSuccess)
        hipLaunchKernelGGL(k_node_capacity, dim3(grid1(ctx->N, 256, 8192)), dim3(256), 0, ctx->stream, egrid(ctx),
                           (const double*)ce, ctx->dcap);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(ce);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(ctx, PF_EDEVICE, "pf_capacity_set: device copy failed");
    const bool entering = !ctx->dyn;
    ctx->theta = theta;
    ctx->dyn = true;
    if ((rc = rebuild_dsh(ctx))) return rc;
    if (ctx->have_coeffs) {
        if (entering) dyn_capture(ctx);  // kh / kv hold the stiffness as set
        else dyn_apply(ctx);             // khB / kvB do; theta may have changed
    }
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->dirty = true;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_dynamic_step(pf_ctx* ctx, const pf_solve_opts* opts, int held, pf_solve_stats* stats) {
    if (!ctx || !opts) return PF_EINVAL;
    if (held != PF_HELD_FIXED && held != PF_HELD_REFERENCE) return fail(ctx, PF_EINVAL, "held must be PF_HELD_*");
    if (!ctx->dyn) return fail(ctx, PF_ESTATE, "pf_capacity_set first");
    if (ctx->have_dconv || ctx->have_bnd) return fail(ctx, PF_ESTATE, "a boundary matrix is set");
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no temperature on the device (pf_solution_set)");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = prepare(ctx))) return rc;
    const long long N = ctx->N;
    const double omt = 1. - ctx->theta;
    // K T^n over all nodes (no elimination), then F = F0 + C T^n - (1 - theta) K T^n
    launch_apply<double, MODE_OP, false, false>(ctx, ctx->x, nullptr, nullptr, nullptr, ctx->tmp, nullptr, ctx->khB,
                                                ctx->kvB, nullptr, true);
    const bool quirk = held == PF_HELD_REFERENCE && ctx->nd;
    if (quirk) {
        if ((size_t)ctx->nd > ctx->ydir_cap) {
            if ((rc = dalloc(ctx, &ctx->ydir, (size_t)ctx->nd))) return rc;
            ctx->ydir_cap = (size_t)ctx->nd;
        }
        hipLaunchKernelGGL(k_dyn_dir, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd, omt,
                           (const long long*)ctx->dnodes, (const double*)ctx->dcap, (const double*)ctx->x,
                           (const double*)ctx->tmp, ctx->ydir);
    }
    hipLaunchKernelGGL(k_dyn_rhs, dim3(grid1(N, 256, 4096)), dim3(256), 0, ctx->stream, N, omt, (const double*)ctx->F0,
                       (const double*)ctx->dcap, (const double*)ctx->x, (const double*)ctx->tmp, ctx->F);
    HIPCK(hipGetLastError());
    pf_solve_opts o = *opts;
    o.x0 = PF_X0_DEVICE;  // from T^n
    rc = pf_solve(ctx, &o, nullptr, stats);
    if (rc != PF_OK && rc != PF_ENOCONV) return rc;
    if (quirk)
        hipLaunchKernelGGL(k_dyn_dir_add, dim3(grid1(ctx->nd)), dim3(256), 0, ctx->stream, (long long)ctx->nd,
                           (const long long*)ctx->dnodes, (const double*)ctx->ydir, ctx->x);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    return rc;
}

int pf_junctions_set(pf_ctx* ctx, int64_t nje, int ncorner, const int64_t* elems, const int64_t* bottom_nodes,
                     const int64_t* top_nodes, const int* junction, int nj, const double* heights) {
    if (!ctx) return PF_EINVAL;
    if (nje < 0 || (nje > 0 && (!elems || !bottom_nodes || !top_nodes || !junction || !heights)) || ncorner < 1 ||
        ncorner > 4)
        return fail(ctx, PF_EINVAL, "bad junction description");
    HIPCK(hipSetDevice(ctx->device));
    for (int64_t q = 0; q < nje; ++q) {
        if (elems[q] < 0 || elems[q] >= ctx->E || junction[q] < 0 || junction[q] >= nj)
            return fail(ctx, PF_EINVAL, "junction element out of range");
        for (int c = 0; c < ncorner; ++c)
            if (bottom_nodes[q * ncorner + c] < 0 || bottom_nodes[q * ncorner + c] >= ctx->N ||
                top_nodes[q * ncorner + c] < 0 || top_nodes[q * ncorner + c] >= ctx->N)
                return fail(ctx, PF_EINVAL, "junction node out of range");
    }
    for (int n = 0; n < nj; ++n)
        if (!(heights[n] > 0.)) return fail(ctx, PF_EINVAL, "junction height must be positive");
    int rc;
    size_t m = std::max<int64_t>(1, nje);
    if ((rc = dalloc(ctx, &ctx->je, m))) return rc;
    if ((rc = dalloc(ctx, &ctx->jbot, m * ncorner))) return rc;
    if ((rc = dalloc(ctx, &ctx->jtop, m * ncorner))) return rc;
    if ((rc = dalloc(ctx, &ctx->jid, m))) return rc;
    if ((rc = dalloc(ctx, &ctx->jheight, std::max(1, nj)))) return rc;
    if ((rc = dalloc(ctx, &ctx->jbeta, m))) return rc;
    if ((rc = dalloc(ctx, &ctx->jjs, m))) return rc;
    if (nje > 0) {
        HIPCK(hipMemcpy(ctx->je, elems, sizeof(long long) * nje, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->jbot, bottom_nodes, sizeof(long long) * nje * ncorner, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->jtop, top_nodes, sizeof(long long) * nje * ncorner, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->jid, junction, sizeof(int) * nje, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->jheight, heights, sizeof(double) * nj, hipMemcpyHostToDevice));
    }
    ctx->nje = nje;
    ctx->jcorner = ncorner;
    return PF_OK;
}

int pf_junction_params(pf_ctx* ctx, const double* beta, const double* js, int64_t nje) {
    if (!ctx || !beta || !js) return PF_EINVAL;
    if (nje != ctx->nje) return fail(ctx, PF_EINVAL, "one beta and js per junction element");
    for (int64_t q = 0; q < nje; ++q)
        if (!(js[q] > 0.) || !std::isfinite(beta[q])) return fail(ctx, PF_EINVAL, "js must be > 0 and beta finite");
    HIPCK(hipSetDevice(ctx->device));
    if (nje > 0) {
        HIPCK(hipMemcpy(ctx->jbeta, beta, sizeof(double) * nje, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->jjs, js, sizeof(double) * nje, hipMemcpyHostToDevice));
    }
    return PF_OK;
}

int pf_junction_update(pf_ctx* ctx, int stable) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no potential on the device");
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->nje > 0) {
        hipLaunchKernelGGL(k_junction_update, dim3(grid1(ctx->nje, 256, 4096)), dim3(256), 0, ctx->stream,
                           (long long)ctx->nje, ctx->jcorner, ctx->je, ctx->jbot, ctx->jtop, ctx->jid, ctx->jheight,
                           ctx->jbeta, ctx->jjs, stable, ctx->x, ctx->kh, ctx->kv, ctx->kh32, ctx->kv32);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(ctx->stream));
        ctx->dirty = true;
    }
    return PF_OK;
}

int pf_junction_state(pf_ctx* ctx, double* U, double* jy, int64_t nje) {
    if (!ctx) return PF_EINVAL;
    if (nje != ctx->nje) return fail(ctx, PF_EINVAL, "one value per junction element");
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no potential on the device");
    if (nje == 0) return PF_OK;
    HIPCK(hipSetDevice(ctx->device));
    double* d = nullptr;
    HIPCK(hipMalloc((void**)&d, sizeof(double) * 2 * nje));
    hipLaunchKernelGGL(k_junction_state, dim3(grid1(nje, 256, 4096)), dim3(256), 0, ctx->stream, (long long)nje,
                       ctx->jcorner, (const long long*)ctx->je, (const long long*)ctx->jbot, (const long long*)ctx->jtop,
                       (const int*)ctx->jid, (const double*)ctx->jheight, (const double*)ctx->x, (const double*)ctx->kv,
                       d, d + nje);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && U) e = hipMemcpyAsync(U, d, sizeof(double) * nje, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && jy) e = hipMemcpyAsync(jy, d + nje, sizeof(double) * nje, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess || e2 != hipSuccess) return fail(ctx, PF_EDEVICE, "pf_junction_state: device error");
    return PF_OK;
}

int pf_junction_cond_set(pf_ctx* ctx, const double* cond_h, const double* cond_v, int64_t nje, int stable) {
    if (!ctx || !cond_h || !cond_v) return PF_EINVAL;
    if (nje != ctx->nje) return fail(ctx, PF_EINVAL, "one value per junction element");
    if (!ctx->have_coeffs) return fail(ctx, PF_ESTATE, "pf_coeffs_set first");
    for (int64_t q = 0; q < nje; ++q)
        if (std::isinf(cond_h[q]) || std::isnan(cond_h[q]) || std::isinf(cond_v[q]))
            return fail(ctx, PF_EINVAL, "junction conductivity not finite (element " + std::to_string(q) + ")");
    if (nje == 0) return PF_OK;
    HIPCK(hipSetDevice(ctx->device));
    double* d = nullptr;
    HIPCK(hipMalloc((void**)&d, sizeof(double) * 2 * nje));
    hipError_t e = hipMemcpyAsync(d, cond_h, sizeof(double) * nje, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + nje, cond_v, sizeof(double) * nje, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_junction_cond_set, dim3(grid1(nje, 256, 4096)), dim3(256), 0, ctx->stream,
                           (long long)nje, (const long long*)ctx->je, (const double*)d, (const double*)(d + nje),
                           stable, ctx->kh, ctx->kv, ctx->kh32, ctx->kv32);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess || e2 != hipSuccess) return fail(ctx, PF_EDEVICE, "pf_junction_cond_set: device error");
    ctx->dirty = true;
    return PF_OK;
}

int pf_element_flags_set(pf_ctx* ctx, const unsigned char* select, const unsigned char* noheat) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if (select) {
        if (!ctx->sel && (rc = dalloc(ctx, &ctx->sel, ctx->E))) return rc;
        HIPCK(hipMemcpy(ctx->sel, select, ctx->E, hipMemcpyHostToDevice));
    }
    if (noheat) {
        if (!ctx->noheat && (rc = dalloc(ctx, &ctx->noheat, ctx->E))) return rc;
        HIPCK(hipMemcpy(ctx->noheat, noheat, ctx->E, hipMemcpyHostToDevice));
    }
    ctx->have_sel = select != nullptr;
    ctx->have_noheat = noheat != nullptr;
    return PF_OK;
}

int pf_currents(pf_ctx* ctx, double* max_dj2, double* max_j2) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no potential on the device");
    HIPCK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_currents, dim3(grid1(ctx->E, 256, 4096)), dim3(256), 0, ctx->stream, egrid(ctx),
                       (long long)ctx->E, ctx->hx, ctx->hy, ctx->hz, ctx->x, ctx->kh, ctx->kv,
                       ctx->have_sel ? ctx->sel : nullptr, ctx->have_noheat ? ctx->noheat : nullptr, ctx->cur,
                       ctx->heat, ctx->dout + 2, ctx->partials, &ctx->st->counter[3]);
    HIPCK(hipGetLastError());
    double h[2];
    HIPCK(hipMemcpyAsync(h, ctx->dout + 2, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (max_dj2) *max_dj2 = h[0];
    if (max_j2) *max_j2 = h[1];
    ctx->have_cur = true;
    return PF_OK;
}

int pf_postprocess(pf_ctx* ctx, int what, const double* k_h, const double* k_v, double* out) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    HIPCK(hipSetDevice(ctx->device));
    const size_t E = (size_t)ctx->E, dim = (size_t)ctx->dim;
    if (what == PF_POST_CURRENT || what == PF_POST_JOULE) {
        if (!ctx->have_cur) return fail(ctx, PF_ESTATE, "no currents: run pf_currents after a solve");
        if (out) {
            if (what == PF_POST_CURRENT)
                HIPCK(hipMemcpy(out, ctx->cur, sizeof(double) * E * dim, hipMemcpyDeviceToHost));
            else
                HIPCK(hipMemcpy(out, ctx->heat, sizeof(double) * E, hipMemcpyDeviceToHost));
        }
        return PF_OK;
    }
    if (what != PF_POST_FLUX) return fail(ctx, PF_EINVAL, "what must be PF_POST_FLUX, PF_POST_CURRENT or PF_POST_JOULE");
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "no temperature on the device");
    if ((k_h == nullptr) != (k_v == nullptr)) return fail(ctx, PF_EINVAL, "give both conductivity arrays or neither");
    if (!k_h && !ctx->have_law && !ctx->have_coeffs) return fail(ctx, PF_ESTATE, "no conductivities");
    int rc;
    if (!ctx->flux && (rc = dalloc(ctx, &ctx->flux, E * dim))) return rc;
    const double *kh = nullptr, *kv = nullptr, *ah = nullptr, *av = nullptr;
    if (k_h) {
        if (!ctx->kpost && (rc = dalloc(ctx, &ctx->kpost, 2 * E))) return rc;
        HIPCK(hipMemcpyAsync(ctx->kpost, k_h, sizeof(double) * E, hipMemcpyHostToDevice, ctx->stream));
        HIPCK(hipMemcpyAsync(ctx->kpost + E, k_v, sizeof(double) * E, hipMemcpyHostToDevice, ctx->stream));
        kh = ctx->kpost;
        kv = ctx->kpost + E;
    } else if (ctx->have_law) {  // the law at the current temperatures
        kh = ctx->k0h;
        kv = ctx->k0v;
        ah = ctx->ah;
        av = ctx->av;
    } else {  // the conductivities as set
        kh = ctx->dyn ? ctx->khB : ctx->kh;
        kv = ctx->dyn ? ctx->kvB : ctx->kv;
    }
    hipLaunchKernelGGL(k_heat_flux, dim3(grid1(ctx->E, 256, 8192)), dim3(256), 0, ctx->stream, egrid(ctx),
                       (long long)ctx->E, (const double*)ctx->hx, (const double*)ctx->hy, (const double*)ctx->hz,
                       (const double*)ctx->x, kh, kv, ah, av, ctx->flux);
    HIPCK(hipGetLastError());
    if (out) HIPCK(hipMemcpyAsync(out, ctx->flux, sizeof(double) * E * dim, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->have_flux = true;
    return PF_OK;
}

int pf_interp_elements(pf_ctx* ctx, int what, const double* pts, int64_t npts, double fill, double* out) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (npts < 0 || (npts > 0 && (!pts || !out))) return fail(ctx, PF_EINVAL, "bad point list");
    const double* f = nullptr;
    int nc = 1;
    if (what == PF_POST_FLUX) {
        if (!ctx->have_flux) return fail(ctx, PF_ESTATE, "no heat flux: pf_postprocess(PF_POST_FLUX) first");
        f = ctx->flux;
        nc = ctx->dim;
    } else if (what == PF_POST_CURRENT || what == PF_POST_JOULE) {
        if (!ctx->have_cur) return fail(ctx, PF_ESTATE, "no currents: run pf_currents after a solve");
        f = what == PF_POST_CURRENT ? ctx->cur : ctx->heat;
        nc = what == PF_POST_CURRENT ? ctx->dim : 1;
    } else {
        return fail(ctx, PF_EINVAL, "what must be PF_POST_FLUX, PF_POST_CURRENT or PF_POST_JOULE");
    }
    if (npts == 0) return PF_OK;
    HIPCK(hipSetDevice(ctx->device));
    const int dim = ctx->dim;
    std::vector<double> axcat, midcat;
    for (int a = 0; a < dim; ++a) {
        const std::vector<double>& A = ctx->ax[a];
        axcat.insert(axcat.end(), A.begin(), A.end());
        for (size_t i = 0; i + 1 < A.size(); ++i) midcat.push_back(0.5 * (A[i] + A[i + 1]));
    }
    int rc;
    if ((size_t)npts > ctx->ipts_cap) {
        if ((rc = dalloc(ctx, &ctx->ipts, (size_t)npts * dim))) return rc;
        ctx->ipts_cap = (size_t)npts;
    }
    if ((size_t)npts * nc > ctx->iout_cap) {
        if ((rc = dalloc(ctx, &ctx->iout, (size_t)npts * nc))) return rc;
        ctx->iout_cap = (size_t)npts * nc;
    }
    if ((rc = dalloc(ctx, &ctx->iax, axcat.size())) || (rc = dalloc(ctx, &ctx->imid, midcat.size()))) return rc;
    HIPCK(hipMemcpyAsync(ctx->iax, axcat.data(), sizeof(double) * axcat.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->imid, midcat.data(), sizeof(double) * midcat.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->ipts, pts, sizeof(double) * (size_t)npts * dim, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_interp_elems, dim3(grid1(npts, 256, 8192)), dim3(256), 0, ctx->stream, (const double*)ctx->iax,
                       (const double*)ctx->imid, dim, (int)ctx->n[0], (int)ctx->n[1], dim == 3 ? (int)ctx->n[2] : 1, f,
                       nc, (long long)npts, (const double*)ctx->ipts, fill, ctx->iout);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, ctx->iout, sizeof(double) * (size_t)npts * nc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_get_currents(pf_ctx* ctx, double* cur, double* heat) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    if (cur) HIPCK(hipMemcpy(cur, ctx->cur, sizeof(double) * ctx->E * ctx->dim, hipMemcpyDeviceToHost));
    if (heat) HIPCK(hipMemcpy(heat, ctx->heat, sizeof(double) * ctx->E, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_set_launch(pf_ctx* ctx, int rows, int planes) {
    if (!ctx) return PF_EINVAL;
    if (!ctx->N) return fail(ctx, PF_ESTATE, "pf_grid_set first");
    if (ctx->dim == 3) {
        if (rows < 2 || rows * ctx->n[0] > 1024 || planes < 1)
            return fail(ctx, PF_EINVAL, "3D: need rows >= 2, rows * n0 <= 1024, planes >= 1");
        if (tile_lds(rows, (int)ctx->n[0]) > 65536)
            return fail(ctx, PF_EINVAL, "3D: rows * n0 too large for the tile kernel's LDS (64 KiB)");
        int ytiles = (int)((ctx->n[1] + rows - 2) / (rows - 1));
        size_t need = 2 * (((size_t)ytiles * ((ctx->n[2] + planes - 1) / planes) + 7) / 8 * 8);
        if (need > ctx->npartials) return fail(ctx, PF_EINVAL, "too many blocks for the reduction buffer");
        ctx->BY[0] = ctx->BY[1] = rows;
        ctx->KC[0] = ctx->KC[1] = planes;
    } else {
        if (planes < 1) return fail(ctx, PF_EINVAL, "planes >= 1");
        size_t need = 2 * (size_t)((ctx->n[1] + planes - 1) / planes);
        if (need > ctx->npartials) return fail(ctx, PF_EINVAL, "too many blocks for the reduction buffer");
        ctx->KC2 = planes;
    }
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_line_segment(pf_ctx* ctx, int seg) {
    if (!ctx) return PF_EINVAL;
    if (seg > 0 && seg != 4 && seg != 8 && seg != 16 && seg != 32)
        return fail(ctx, PF_EINVAL, "partition length must be 4, 8, 16 or 32 (0: default, < 0: sequential)");
    ctx->line_m = seg;
    if (!ctx->N) return PF_OK;
    return line_config(ctx);
}

int pf_set_multigrid(pf_ctx* ctx, int max_levels, int coarse_sweeps, double omega) {
    if (!ctx) return PF_EINVAL;
    if (max_levels < 0 || max_levels > 32) return fail(ctx, PF_EINVAL, "max_levels must be 0 (no limit) to 32");
    if (coarse_sweeps < 1 || coarse_sweeps > 64) return fail(ctx, PF_EINVAL, "coarse_sweeps must be 1 to 64");
    if (!(omega > 0. && omega <= 0.66)) return fail(ctx, PF_EINVAL, "omega must be in (0, 0.66]");
    ctx->mg_max_levels = max_levels;
    ctx->mg_coarse_sweeps = coarse_sweeps;
    ctx->mg_omega = omega;
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->stream) HIPCK(hipStreamSynchronize(ctx->stream));
    mg_free(ctx);  // rebuilt by the next multigrid solve
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_multigrid_xcd(pf_ctx* ctx, int mode) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->stream) HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->mg_xcd = mode < 0 ? 0 : (mode > 2 ? 2 : mode);
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_multigrid_fusion(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    if (ctx->stream) HIPCK(hipStreamSynchronize(ctx->stream));
    ctx->mg_fuse = on ? 1 : 0;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_multigrid_levels(const pf_ctx* ctx) { return ctx && ctx->mg_built ? (int)ctx->mg.size() + 1 : 0; }

int pf_set_apply(pf_ctx* ctx, int precision, int kind, int planes) {
    if (!ctx) return PF_EINVAL;
    if (kind != 0 && kind != 1) return fail(ctx, PF_EINVAL, "apply kind 0 (tile blocks) or 1 (row blocks)");
    if (precision != PF_FP64 && precision != PF_FP32_REFINE && precision != PF_FP32_RELIABLE && precision != -1)
        return fail(ctx, PF_EINVAL, "precision PF_FP64, PF_FP32_REFINE or -1 (both)");
    if (kind == 0 && ctx->dim == 3 && tile_lds(ctx->BY[0], (int)ctx->n[0]) > 65536)
        return fail(ctx, PF_EINVAL, "tile blocks: rows * n0 too large for the kernel's LDS (64 KiB)");
    if (kind == 1) {
        if (ctx->dim == 3 && ctx->n[0] > 1024) return fail(ctx, PF_EINVAL, "row blocks need n0 <= 1024");
        if (planes < 4) return fail(ctx, PF_EINVAL, "row blocks: planes >= 4");
    }
    for (int pk = 0; pk < 2; ++pk) {
        if (precision != -1 && pk != (precision == PF_FP64 ? 0 : 1)) continue;
        ctx->apply_kind[pk] = kind;
        if (kind == 1) ctx->KR[pk] = planes;
    }
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_block_order(pf_ctx* ctx, int xcd) {
    if (!ctx) return PF_EINVAL;
    ctx->xcd_order = xcd != 0;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_graphs(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    ctx->graphs = on != 0;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_graphs_enabled(const pf_ctx* ctx) { return ctx && ctx->graphs ? 1 : 0; }

int pf_set_prefetch(pf_ctx* ctx, int depth) {
    if (!ctx) return PF_EINVAL;
    if (depth != 1 && depth != 2) return fail(ctx, PF_EINVAL, "prefetch depth 1 or 2");
    ctx->pfd = depth;
    drop_graphs(ctx);
    return PF_OK;
}

int pf_set_profile(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    ctx->profile = on != 0;
    return PF_OK;
}

// Device copy b = a, the achievable-bandwidth probe and the calibration kernel for the
// hardware traffic counters: one 16-byte element per thread over a grid that covers the
// array, so the blocks stream through memory in address order (measured on one MI355X for
// 1 GiB: 6.25 TB/s, against 5.6 TB/s for a grid-stride loop over 4 blocks per CU and
// 4.2-4.9 TB/s with several loads per thread spaced a grid apart).
typedef double pf_d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_stream_copy(long long n16, const pf_d2* __restrict__ a,
                                                     pf_d2* __restrict__ b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) b[i] = a[i];
}

int pf_stream_copy(pf_ctx* ctx, int64_t bytes, int reps, double* seconds) {
    if (!ctx || bytes < 16 || reps <= 0 || !seconds) return PF_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    const long long n16 = bytes / 16;
    pf_d2 *a = nullptr, *b = nullptr;
    HIPCK(hipMalloc((void**)&a, 16 * (size_t)n16));
    if (hipMalloc((void**)&b, 16 * (size_t)n16) != hipSuccess) {
        hipFree(a);
        return fail(ctx, PF_EDEVICE, "pf_stream_copy: out of device memory");
    }
    hipMemsetAsync(a, 0, 16 * (size_t)n16, ctx->stream);
    if ((n16 + 255) / 256 > 0x7fffffffll) return fail(ctx, PF_EINVAL, "pf_stream_copy: size too large");
    const int blocks = (int)((n16 + 255) / 256);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ctx->stream, n16, a, b);
    hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ctx->stream, n16, a, b);
    hipEventRecord(e1, ctx->stream);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (err != hipSuccess) return fail(ctx, PF_EDEVICE, std::string("pf_stream_copy: ") + hipGetErrorString(err));
    *seconds = 1e-3 * ms / reps;
    return PF_OK;
}

}  // extern "C"

// Back-to-back launches of the two kernels of one PCG iteration on the state the
// last solve of this precision left (as a kernel trace sees them inside a solve):
// the fused operator apply (p = z + beta p_old, q = A p, (p, q), x += alpha p_old)
// and the preconditioner update (r -= alpha q, z = M^-1 r, (r, z), (r, r)).  The
// iteration is not advanced in any way that matters: x is saved and restored.
template <typename T>
static int time_kernels(pf_ctx* ctx, int precond, int reps, double* t_apply, double* t_update, T* x, T* r, T* p,
                        T* p2, T* q, T* z, const T* dinv, const T* la, const T* lw, const T* lcp, const T* kh,
                        const T* kv) {
    const long long N = ctx->N;
    T* save = reinterpret_cast<T*>(ctx->tmp);  // N doubles of scratch hold N values of T
    HIPCK(hipMemcpyAsync(save, x, sizeof(T) * N, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    PcgState s = *ctx->hst;
    s.done = 0;
    s.status = 0;
    s.iters = std::max(s.iters, 1);
    s.maxit = 1 << 30;
    s.stop2 = 0.;
    if (!(s.rz_prev > 0.)) s.rz_prev = 1.;
    if (!(s.rz > 0.)) s.rz = 1.;
    *ctx->hst = s;
    HIPCK(hipMemcpyAsync(ctx->st, ctx->hst, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    const int ncol = (int)(ctx->dim == 3 ? ctx->n[0] * ctx->n[1] : ctx->n[0]);
    const int nplanes = (int)(ctx->dim == 3 ? ctx->n[2] : ctx->n[1]);
    const int pm = sizeof(T) == 8 ? ctx->pm64 : ctx->pm32;
    const T* schur = (const T*)(sizeof(T) == 8 ? (const void*)ctx->schur : (const void*)ctx->schur32);
    PartShape ps = part_shape<T>(pm, nplanes, ncol);
    hipEvent_t e[3];
    for (auto& ev : e) HIPCK(hipEventCreate(&ev));
    launch_apply<T, MODE_OP, true, true>(ctx, nullptr, z, p, p2, q, nullptr, kh, kv, x);  // warm-up
    hipEventRecord(e[0], ctx->stream);
    for (int i = 0; i < reps; ++i) launch_apply<T, MODE_OP, true, true>(ctx, nullptr, z, p, p2, q, nullptr, kh, kv, x);
    hipEventRecord(e[1], ctx->stream);
    for (int i = 0; i < reps; ++i) {
        if (precond == 2 && pm)
            launch_part<T, false>(ctx, ps, ncol, nplanes, ncol, r, q, la, lw, lcp, schur, z);
        else if (precond == 2)
            pf_launch(ctx, (k_cg_line<T, false>), dim3((ncol + 63) / 64, 1), dim3(64), 0, ctx->stream, ncol,
                               nplanes, (long long)ncol, nplanes, r, (const T*)q, la, lw, lcp, z, ctx->st,
                               ctx->partials);
        else
            pf_launch(ctx, (k_cg_update<T>), dim3(ctx->upd_blocks), dim3(256), 0, ctx->stream, N, r, q, dinv, z,
                               ctx->st, ctx->partials);
    }
    hipEventRecord(e[2], ctx->stream);
    HIPCK(hipGetLastError());
    HIPCK(hipEventSynchronize(e[2]));
    float m1 = 0.f, m2 = 0.f;
    hipEventElapsedTime(&m1, e[0], e[1]);
    hipEventElapsedTime(&m2, e[1], e[2]);
    for (auto& ev : e) hipEventDestroy(ev);
    *t_apply = 1e-3 * m1 / reps;
    *t_update = 1e-3 * m2 / reps;
    HIPCK(hipMemcpyAsync(x, save, sizeof(T) * N, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->hst->done = 1;
    HIPCK(hipMemcpyAsync(&ctx->st->done, &ctx->hst->done, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

extern "C" {

int pf_time_kernels(pf_ctx* ctx, int precision, int precond, int reps, double* t_apply, double* t_update) {
    if (!ctx || reps <= 0 || !t_apply || !t_update) return PF_EINVAL;
    if (!ctx->have_solution) return fail(ctx, PF_ESTATE, "pf_time_kernels: run a solve of this precision first");
    HIPCK(hipSetDevice(ctx->device));
    const int pc = precond == 1 ? 1 : 2;
    if (precision == PF_FP64)
        return time_kernels<double>(ctx, pc, reps, t_apply, t_update, ctx->x, ctx->r, ctx->p, ctx->p2, ctx->qv,
                                    ctx->z, ctx->dinv, ctx->la, ctx->lw, ctx->lcp, ctx->kh, ctx->kv);
    return time_kernels<float>(ctx, pc, reps, t_apply, t_update, ctx->x32, ctx->r32, ctx->p32, ctx->p2_32, ctx->q32,
                               ctx->z32, ctx->dinv32, ctx->la32, ctx->lw32, ctx->lcp32, ctx->kh32, ctx->kv32);
}

int pf_time_apply(pf_ctx* ctx, int precision, int reps, double* seconds) {
    if (!ctx || reps <= 0 || !seconds) return PF_EINVAL;
    if (!ctx->have_coeffs) return fail(ctx, PF_ESTATE, "pf_coeffs_set first");
    HIPCK(hipSetDevice(ctx->device));
    hipEvent_t a, b;
    HIPCK(hipEventCreate(&a));
    HIPCK(hipEventCreate(&b));
    auto go = [&]() {
        if (precision == PF_FP64)
            launch_apply<double, MODE_OP, false, false>(ctx, ctx->p, nullptr, nullptr, nullptr, ctx->qv, nullptr, ctx->kh, ctx->kv);
        else
            launch_apply<float, MODE_OP, false, false>(ctx, ctx->p32, nullptr, nullptr, nullptr, ctx->q32, nullptr, ctx->kh32,
                                                       ctx->kv32);
    };
    go();
    HIPCK(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps; ++i) go();
    HIPCK(hipEventRecord(b, ctx->stream));
    HIPCK(hipGetLastError());
    HIPCK(hipEventSynchronize(b));
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    *seconds = 1e-3 * ms / reps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return PF_OK;
}

}  // extern "C"

// ===========================================================================
// Batched 2D Shockley solver (persistent kernel over a work queue)
// ===========================================================================
struct pf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int ninst = 0;
    long long N = 0, E = 0;
    BatchInst* dinst = nullptr;
    int* dorder = nullptr;
    unsigned int* dqueue = nullptr;
    double *h0 = nullptr, *h1 = nullptr, *rm = nullptr;
    double *kh = nullptr, *kv = nullptr, *kh0 = nullptr, *kv0 = nullptr;
    uint64_t* dmask = nullptr;
    double *x = nullptr, *x0 = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *la = nullptr,
           *lw = nullptr, *lc = nullptr, *cur = nullptr;
    BatchStats* dstats = nullptr;
    BatchResume* dres = nullptr;
    unsigned int* hfin = nullptr;  // pinned: instances finished, read between launches
    // multigrid levels of all instances (pools described by BatchLevel)
    BatchMg mg{};
    std::vector<long long> node_off;
    int tuned_block = 0, tuned_nwg = 0;
    std::vector<void*> allocs;
    hipGraphExec_t graph = nullptr, graph_init = nullptr;
    bool graphs = true;
    int graph_block = 0, graph_nwg = 0;
    BatchParams graph_prm{};
};

#define BHIPCK(call)                                                                \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            b->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return PF_EDEVICE;                                                      \
        }                                                                           \
    } while (0)

template <typename T> static int balloc(pf_batch* b, T** ptr, size_t count) {
    BHIPCK(hipMalloc((void**)ptr, sizeof(T) * std::max<size_t>(count, 1)));
    b->allocs.push_back((void*)*ptr);
    return PF_OK;
}

__global__ void k_copy_d(long long n, const double* __restrict__ a, double* __restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

extern "C" {

int pf_batch_create(pf_batch** out, int device) {
    if (!out) return PF_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PF_EDEVICE;
    if (device < 0 || device >= ndev) return PF_EINVAL;
    pf_batch* b = new pf_batch();
    b->device = device;
    b->graphs = graphs_default("plask_amd batch");
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return PF_EDEVICE;
    }
    if (hipHostMalloc((void**)&b->hfin, sizeof(unsigned int), hipHostMallocDefault) != hipSuccess) {
        hipStreamDestroy(b->stream);
        delete b;
        return PF_EDEVICE;
    }
    *out = b;
    return PF_OK;
}

void pf_batch_destroy(pf_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    hipStreamSynchronize(b->stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    if (b->graph_init) hipGraphExecDestroy(b->graph_init);
    for (void* p : b->allocs) hipFree(p);
    if (b->hfin) hipHostFree(b->hfin);
    hipStreamDestroy(b->stream);
    delete b;
}

const char* pf_batch_last_error(const pf_batch* b) { return b ? b->err.c_str() : "null batch"; }

int pf_batch_set(pf_batch* b, int ninst, int cylindrical, const int32_t* n0, const int32_t* n1, const double* ax0,
                 const double* ax1, const double* sh, const double* sv, const int64_t* ndir, const int64_t* dnodes,
                 const double* dvals, const int32_t* jlo, const int32_t* jhi, const double* beta, const double* js) {
    if (!b || ninst <= 0 || !n0 || !n1 || !ax0 || !ax1 || !sh || !sv || !ndir || !jlo || !jhi || !beta || !js)
        return PF_EINVAL;
    if (b->ninst) {
        b->err = "pf_batch_set may be called once per batch object";
        return PF_ESTATE;
    }
    BHIPCK(hipSetDevice(b->device));
    std::vector<BatchInst> inst(ninst);
    std::vector<double> h0, h1, rm;
    long long N = 0, E = 0, a0 = 0, a1 = 0, dpos = 0;
    b->node_off.resize(ninst);
    for (int i = 0; i < ninst; ++i) {
        if (n0[i] < 2 || n1[i] < 2) {
            b->err = "instance " + std::to_string(i) + ": each axis needs at least 2 nodes";
            return PF_EINVAL;
        }
        BatchInst& I = inst[i];
        I.n0 = n0[i];
        I.n1 = n1[i];
        I.node_off = N;
        I.elem_off = E;
        I.h0_off = (int)h0.size();
        I.h1_off = (int)h1.size();
        for (int k = 0; k + 1 < n0[i]; ++k) {
            double w = ax0[a0 + k + 1] - ax0[a0 + k];
            if (!(w > 0.)) {
                b->err = "instance " + std::to_string(i) + ": axis 0 must increase";
                return PF_EINVAL;
            }
            h0.push_back(w);
            rm.push_back(0.5 * (ax0[a0 + k] + ax0[a0 + k + 1]));
        }
        for (int k = 0; k + 1 < n1[i]; ++k) {
            double h = ax1[a1 + k + 1] - ax1[a1 + k];
            if (!(h > 0.)) {
                b->err = "instance " + std::to_string(i) + ": axis 1 must increase";
                return PF_EINVAL;
            }
            h1.push_back(h);
        }
        if (jlo[i] < 0 || jhi[i] < jlo[i] || jhi[i] > n1[i] - 1 || !(js[i] > 0.)) {
            b->err = "instance " + std::to_string(i) + ": bad junction rows or js";
            return PF_EINVAL;
        }
        I.jlo = jlo[i];
        I.jhi = jhi[i];
        I.jheight = ax1[a1 + jhi[i]] - ax1[a1 + jlo[i]];
        I.beta = beta[i];
        I.js = js[i];
        I.cyl = cylindrical;
        b->node_off[i] = N;
        a0 += n0[i];
        a1 += n1[i];
        N += (long long)n0[i] * n1[i];
        E += (long long)(n0[i] - 1) * (n1[i] - 1);
    }
    std::vector<double> x0(N, 0.);
    std::vector<uint64_t> mask((N + 63) / 64, 0ull);
    for (int i = 0; i < ninst; ++i) {
        long long Ni = (long long)n0[i] * n1[i];
        for (int64_t d = 0; d < ndir[i]; ++d) {
            int64_t nd = dnodes[dpos + d];
            if (nd < 0 || nd >= Ni) {
                b->err = "instance " + std::to_string(i) + ": Dirichlet node out of range";
                return PF_EINVAL;
            }
            long long g = inst[i].node_off + nd;
            mask[g >> 6] |= 1ull << (g & 63);
            x0[g] = dvals[dpos + d];
        }
        dpos += ndir[i];
    }
    // multigrid levels per instance: axis 0 coarsened while it has more than 3 nodes, all
    // rows of axis 1 kept; the Dirichlet set is fixed, so the coarse masks are made here
    std::vector<BatchLevel> levs;
    std::vector<int> tcx, tmx;
    std::vector<double> twx, lh0, lrm;
    std::vector<uint64_t> lmask;
    long long lnode = 0, lelem = 0, lbits = 0;
    {
        long long ax0_pos = 0;
        for (int i = 0; i < ninst; ++i) {
            BatchInst& I = inst[i];
            const int m1 = n1[i];
            std::vector<double> f(ax0 + ax0_pos, ax0 + ax0_pos + n0[i]);
            ax0_pos += n0[i];
            std::vector<char> fm((size_t)n0[i] * m1, 0);
            for (long long nd = 0; nd < (long long)n0[i] * m1; ++nd) {
                const long long g = I.node_off + nd;
                fm[nd] = (char)((mask[g >> 6] >> (g & 63)) & 1ull);
            }
            I.lev_off = (int)levs.size();
            I.nlev = 0;
            while (f.size() > 3) {
                std::vector<int> ci, mm;
                std::vector<double> ww, cc;
                mg_coarsen_axis(f, true, ci, mm, ww, cc);
                BatchLevel B{};
                B.n0 = (int)cc.size();
                B.cx_off = (int)tcx.size();
                tcx.insert(tcx.end(), ci.begin(), ci.end());
                B.mx_off = (int)tmx.size();
                tmx.insert(tmx.end(), mm.begin(), mm.end());
                twx.insert(twx.end(), ww.begin(), ww.end());
                B.h0_off = (int)lh0.size();
                for (int k = 0; k + 1 < B.n0; ++k) {
                    lh0.push_back(cc[k + 1] - cc[k]);
                    lrm.push_back(0.5 * (cc[k] + cc[k + 1]));
                }
                B.node_off = lnode;
                B.elem_off = lelem;
                B.mask_off = lbits;
                const int nf = (int)f.size();
                std::vector<char> cm((size_t)B.n0 * m1, 0);
                for (int k = 0; k < m1; ++k)
                    for (int c = 0; c < B.n0; ++c) cm[(size_t)k * B.n0 + c] = fm[(size_t)k * nf + ci[c]];
                lmask.resize((size_t)((lbits + (long long)B.n0 * m1 + 63) / 64), 0ull);
                for (long long nd = 0; nd < (long long)B.n0 * m1; ++nd)
                    if (cm[nd]) lmask[(lbits + nd) >> 6] |= 1ull << ((lbits + nd) & 63);
                lnode += (long long)B.n0 * m1;
                lelem += (long long)(B.n0 - 1) * (m1 - 1);
                lbits += (long long)B.n0 * m1;
                levs.push_back(B);
                I.nlev++;
                f = cc;
                fm.swap(cm);
            }
        }
    }
    // largest instances first (longest-processing-time order for the work queue)
    std::vector<int> order(ninst);
    for (int i = 0; i < ninst; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return (long long)n0[a] * n1[a] > (long long)n0[c] * n1[c];
    });
    b->ninst = ninst;
    b->N = N;
    b->E = E;
    int rc;
    if ((rc = balloc(b, &b->dinst, ninst)) || (rc = balloc(b, &b->dorder, ninst)) || (rc = balloc(b, &b->dqueue, 2)) ||
        (rc = balloc(b, &b->dres, ninst)) ||
        (rc = balloc(b, &b->h0, h0.size())) || (rc = balloc(b, &b->h1, h1.size())) || (rc = balloc(b, &b->rm, rm.size())) ||
        (rc = balloc(b, &b->kh, E)) || (rc = balloc(b, &b->kv, E)) || (rc = balloc(b, &b->kh0, E)) ||
        (rc = balloc(b, &b->kv0, E)) || (rc = balloc(b, &b->dmask, mask.size())) || (rc = balloc(b, &b->x, N)) ||
        (rc = balloc(b, &b->x0, N)) || (rc = balloc(b, &b->r, N)) || (rc = balloc(b, &b->p, N)) ||
        (rc = balloc(b, &b->q, N)) || (rc = balloc(b, &b->z, N)) || (rc = balloc(b, &b->la, N)) ||
        (rc = balloc(b, &b->lw, N)) || (rc = balloc(b, &b->lc, N)) || (rc = balloc(b, &b->cur, 2 * E)) ||
        (rc = balloc(b, &b->dstats, ninst)))
        return rc;
    {
        BatchLevel* dlev = nullptr;
        int *dcx = nullptr, *dmx = nullptr;
        double *dwx = nullptr, *dh0 = nullptr, *drm = nullptr;
        uint64_t* dlm = nullptr;
        BatchMg& M = b->mg;
        if ((rc = balloc(b, &dlev, levs.size())) || (rc = balloc(b, &dcx, tcx.size())) ||
            (rc = balloc(b, &dmx, tmx.size())) || (rc = balloc(b, &dwx, twx.size())) ||
            (rc = balloc(b, &dh0, lh0.size())) || (rc = balloc(b, &drm, lrm.size())) ||
            (rc = balloc(b, &dlm, lmask.size())) || (rc = balloc(b, &M.kh, lelem)) || (rc = balloc(b, &M.kv, lelem)) ||
            (rc = balloc(b, &M.b, lnode)) || (rc = balloc(b, &M.x, lnode)) || (rc = balloc(b, &M.r, lnode)) ||
            (rc = balloc(b, &M.la, lnode)) || (rc = balloc(b, &M.lw, lnode)) || (rc = balloc(b, &M.lc, lnode)))
            return rc;
        if (!levs.empty()) BHIPCK(hipMemcpy(dlev, levs.data(), sizeof(BatchLevel) * levs.size(), hipMemcpyHostToDevice));
        if (!tcx.empty()) BHIPCK(hipMemcpy(dcx, tcx.data(), sizeof(int) * tcx.size(), hipMemcpyHostToDevice));
        if (!tmx.empty()) BHIPCK(hipMemcpy(dmx, tmx.data(), sizeof(int) * tmx.size(), hipMemcpyHostToDevice));
        if (!twx.empty()) BHIPCK(hipMemcpy(dwx, twx.data(), sizeof(double) * twx.size(), hipMemcpyHostToDevice));
        if (!lh0.empty()) BHIPCK(hipMemcpy(dh0, lh0.data(), sizeof(double) * lh0.size(), hipMemcpyHostToDevice));
        if (!lrm.empty()) BHIPCK(hipMemcpy(drm, lrm.data(), sizeof(double) * lrm.size(), hipMemcpyHostToDevice));
        if (!lmask.empty()) BHIPCK(hipMemcpy(dlm, lmask.data(), sizeof(uint64_t) * lmask.size(), hipMemcpyHostToDevice));
        M.lev = dlev;
        M.cx = dcx;
        M.mx = dmx;
        M.wx = dwx;
        M.h0 = dh0;
        M.rm = drm;
        M.mask = dlm;
    }
    BHIPCK(hipMemcpy(b->dinst, inst.data(), sizeof(BatchInst) * ninst, hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->dorder, order.data(), sizeof(int) * ninst, hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->h0, h0.data(), sizeof(double) * h0.size(), hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->h1, h1.data(), sizeof(double) * h1.size(), hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->rm, rm.data(), sizeof(double) * rm.size(), hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->kh0, sh, sizeof(double) * E, hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->kv0, sv, sizeof(double) * E, hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->dmask, mask.data(), sizeof(uint64_t) * mask.size(), hipMemcpyHostToDevice));
    BHIPCK(hipMemcpy(b->x0, x0.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    return PF_OK;
}

}  // extern "C"

// Start of a batch solve: potentials from the Dirichlet values, material
// conductivities, zero currents, every instance not started.
static void batch_launch_init(pf_batch* b) {
    const int gb = 2048;
    hipLaunchKernelGGL(k_copy_d, dim3(gb), dim3(256), 0, b->stream, b->N, b->x0, b->x);
    hipLaunchKernelGGL(k_copy_d, dim3(gb), dim3(256), 0, b->stream, b->E, b->kh0, b->kh);
    hipLaunchKernelGGL(k_copy_d, dim3(gb), dim3(256), 0, b->stream, b->E, b->kv0, b->kv);
    hipLaunchKernelGGL((k_fill<double>), dim3(gb), dim3(256), 0, b->stream, 2 * b->E, b->cur, 0.);
    hipLaunchKernelGGL(k_batch_reset, dim3((b->ninst + 255) / 256), dim3(256), 0, b->stream, b->ninst, b->dstats,
                       b->dres, b->dqueue);
}

// One launch of the persistent kernel: every unfinished instance advances by at most
// prm.loop_budget Picard loops (all of them when the budget is 0).
static void batch_launch_round(pf_batch* b, int block, int nwg, const BatchParams& prm) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, b->stream, b->dqueue);
#define PF_BATCH_LAUNCH(BS)                                                                                         \
    hipLaunchKernelGGL((k_batch_shockley2d<BS>), dim3(nwg), dim3(BS), 0, b->stream, b->dinst, b->dorder, b->ninst, \
                       b->dqueue, prm, b->h0, b->h1, b->rm, b->kh, b->kv, b->dmask, b->x, b->r, b->p, b->q, b->z,   \
                       b->la, b->lw, b->lc, b->cur, b->dstats, b->dres, b->mg)
    if (block == 256) PF_BATCH_LAUNCH(256);
    else if (block == 512) PF_BATCH_LAUNCH(512);
    else PF_BATCH_LAUNCH(1024);
#undef PF_BATCH_LAUNCH
}

static bool same_prm(const BatchParams& a, const BatchParams& c) {
    return a.maxloops == c.maxloops && a.maxit == c.maxit && a.maxerr == c.maxerr && a.rtol == c.rtol &&
           a.stable == c.stable && a.precond == c.precond && a.omega == c.omega && a.coarse_sweeps == c.coarse_sweeps &&
           a.loop_budget == c.loop_budget;
}

// Kernel parameters from the caller's options (0 = the documented default).
static BatchParams batch_params(const pf_batch_opts* o) {
    return BatchParams{o->maxloops > 0 ? o->maxloops : 30, o->maxit > 0 ? o->maxit : 100000,
                       o->maxerr > 0. ? o->maxerr : 0.05, o->rtol > 0. ? o->rtol : 1e-8, o->stable,
                       o->precond == 2 ? 2 : 3, 0.65, 2, o->loop_budget};
}

template <typename F> static int batch_capture(pf_batch* b, hipGraphExec_t* exec, F&& launches) {
    hipGraph_t g;
    BHIPCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launches();
    BHIPCK(hipStreamEndCapture(b->stream, &g));
    BHIPCK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    return PF_OK;
}

// The whole batch: the start launch, then launches of the persistent kernel until every
// instance has finished its Picard loop.  Each launch runs at most prm.loop_budget loops
// per instance, so it returns after a bounded amount of work; between launches the host
// reads the number of finished instances and reports it through `progress`.  Every
// launch advances each unfinished instance by at least one loop, so at most
// ceil(maxloops / loop_budget) launches run.
static int batch_run(pf_batch* b, int block, int nwg, const BatchParams& prm, double* seconds,
                     void (*progress)(int, int, int, double, void*), void* user) {
    int rc;
    if (b->graphs) {
        if (!b->graph_init && (rc = batch_capture(b, &b->graph_init, [&] { batch_launch_init(b); }))) return rc;
        if (!b->graph || b->graph_block != block || b->graph_nwg != nwg || !same_prm(b->graph_prm, prm)) {
            if (b->graph) hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            if ((rc = batch_capture(b, &b->graph, [&] { batch_launch_round(b, block, nwg, prm); }))) return rc;
            b->graph_block = block;
            b->graph_nwg = nwg;
            b->graph_prm = prm;
        }
    }
    const int rounds_max = prm.loop_budget > 0 ? (prm.maxloops + prm.loop_budget - 1) / prm.loop_budget : 1;
    hipEvent_t e0, e1;
    BHIPCK(hipEventCreate(&e0));
    BHIPCK(hipEventCreate(&e1));
    BHIPCK(hipEventRecord(e0, b->stream));
    if (b->graphs) BHIPCK(hipGraphLaunch(b->graph_init, b->stream));
    else batch_launch_init(b);
    auto t0 = std::chrono::steady_clock::now();
    int finished = 0, round = 0;
    while (finished < b->ninst && round < rounds_max) {
        if (b->graphs) BHIPCK(hipGraphLaunch(b->graph, b->stream));
        else batch_launch_round(b, block, nwg, prm);
        BHIPCK(hipMemcpyAsync(b->hfin, b->dqueue + 1, sizeof(unsigned int), hipMemcpyDeviceToHost, b->stream));
        BHIPCK(hipStreamSynchronize(b->stream));
        finished = (int)*b->hfin;
        ++round;
        if (progress)
            progress(round, finished, b->ninst, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                     user);
    }
    BHIPCK(hipEventRecord(e1, b->stream));
    BHIPCK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (seconds) *seconds = 1e-3 * ms;
    if (finished < b->ninst) {
        b->err = "batch: " + std::to_string(b->ninst - finished) + " instances unfinished after " +
                 std::to_string(round) + " launches";
        return PF_ESTATE;
    }
    return PF_OK;
}

extern "C" {

int pf_batch_autotune(pf_batch* b, const pf_batch_opts* o, int* best_block, int* best_nwg, double* best_seconds) {
    if (!b || !o) return PF_EINVAL;
    if (!b->ninst) {
        b->err = "pf_batch_set first";
        return PF_ESTATE;
    }
    BHIPCK(hipSetDevice(b->device));
    if (o->maxloops < 0 || o->loop_budget < 0) {
        b->err = "maxloops and loop_budget must be >= 0";
        return PF_EINVAL;
    }
    const BatchParams prm = batch_params(o);
    int props_cu = 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, b->device) == hipSuccess) props_cu = pr.multiProcessorCount;
    const int blocks[3] = {256, 512, 1024};
    const int mult[2] = {1, 2};
    double best = 1e300;
    int bb = 1024, bn = props_cu;
    for (int bi = 0; bi < 3; ++bi)
        for (int mi = 0; mi < 2; ++mi) {
            int nwg = std::min(b->ninst, props_cu * mult[mi]);
            double t = 0.;
            int rc = batch_run(b, blocks[bi], nwg, prm, &t, o->progress, o->progress_user);
            if (rc) return rc;
            if (t < best) {
                best = t;
                bb = blocks[bi];
                bn = nwg;
            }
        }
    b->tuned_block = bb;
    b->tuned_nwg = bn;
    if (best_block) *best_block = bb;
    if (best_nwg) *best_nwg = bn;
    if (best_seconds) *best_seconds = best;
    return PF_OK;
}

int pf_batch_conductivities(pf_batch* b, double* sh, double* sv) {
    if (!b) return PF_EINVAL;
    if (!b->ninst) {
        b->err = "pf_batch_set first";
        return PF_ESTATE;
    }
    BHIPCK(hipSetDevice(b->device));
    if (sh) BHIPCK(hipMemcpy(sh, b->kh, sizeof(double) * b->E, hipMemcpyDeviceToHost));
    if (sv) BHIPCK(hipMemcpy(sv, b->kv, sizeof(double) * b->E, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_batch_solve(pf_batch* b, const pf_batch_opts* o, double* V_out, pf_batch_stats* stats, double* seconds) {
    if (!b || !o) return PF_EINVAL;
    if (!b->ninst) {
        b->err = "pf_batch_set first";
        return PF_ESTATE;
    }
    BHIPCK(hipSetDevice(b->device));
    int props_cu = 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, b->device) == hipSuccess) props_cu = pr.multiProcessorCount;
    int block = o->block ? o->block : (b->tuned_block ? b->tuned_block : 1024);
    if (block != 256 && block != 512 && block != 1024) {
        b->err = "block must be 256, 512 or 1024";
        return PF_EINVAL;
    }
    int nwg = o->nwg ? o->nwg : (b->tuned_nwg ? b->tuned_nwg : std::min(b->ninst, props_cu));
    if (o->maxloops < 0 || o->loop_budget < 0) {
        b->err = "maxloops and loop_budget must be >= 0";
        return PF_EINVAL;
    }
    const BatchParams prm = batch_params(o);
    int rc = batch_run(b, block, nwg, prm, seconds, o->progress, o->progress_user);
    if (rc) return rc;
    if (V_out) BHIPCK(hipMemcpy(V_out, b->x, sizeof(double) * b->N, hipMemcpyDeviceToHost));
    if (stats) {
        static_assert(sizeof(BatchStats) == sizeof(pf_batch_stats), "stats layout");
        BHIPCK(hipMemcpy(stats, b->dstats, sizeof(BatchStats) * b->ninst, hipMemcpyDeviceToHost));
    }
    return PF_OK;
}

}  // extern "C"
