// gfx950 kernels for the matrix-free Q1 FEM operator and the fused PCG steps.
//
// Operator.  For an axis-aligned box element with diagonal conductivity the Q1
// stiffness matrix is  K = kx' S(x)M(y)M(z) + ky' M(x)S(y)M(z) + kz' M(x)M(y)S(z)
// with S = [1 -1; -1 1], M = [1/3 1/6; 1/6 1/3] and k' the conductivity scaled by
// the element sizes (the closed forms of solvers/thermal/static/therm3d.cpp:218-237
// and therm2d.cpp:216-228).  In the per-dimension basis (sum, difference)
//   sigma = x0 + x1,  delta = x1 - x0
// both S and M are diagonal (S: sigma->0, delta->2 delta; M: sigma->sigma/2,
// delta->delta/6), so K x = G^T diag(L) G x with G the 2^d-point sum/difference
// transform.  One element costs ~55 flops for all 8 outputs instead of 64 FMAs,
// and the two in-plane transform stages are shared by the two element layers that
// touch a node plane.
//
// Data layout in HBM: nodes and elements in natural order, axis 0 fastest
// (node = i + n0*(j + n1*k)).  Element data: two conductivities per element
// (in-plane, vertical) and nothing else; element sizes come from 1D axis arrays.
// Dirichlet nodes are a bitmask (1 bit per node).
//
// Kernel shape (3D): a block holds the whole x extent (BX = n0 threads per row)
// and BY rows in y; thread (tx,ty) owns element column (tx, jb-1+ty) and node
// column (tx, jb-1+ty).  Row ty = 0 is a halo row (its elements feed row 1; its
// nodes are not written).  The block marches along z over a chunk of planes:
// per plane one coalesced load of x into LDS, one LDS exchange of the element
// corner outputs, two barriers.  Nothing is recomputed across z except one halo
// element layer per chunk.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

// MODE_RESG: MODE_RES, skipped after the stop; MODE_RESP: MODE_RESG of x + P x_c, the
// next coarser level's correction interpolated where x is read (MgProl)
enum { MODE_OP = 0, MODE_RES = 1, MODE_RESG = 2, MODE_RESP = 3 };

// Coarse-grid correction of a multigrid level read together with x: x + P x_c at the
// free nodes, P the Q1 interpolation in the plane of the next coarser level (node i of
// an in-plane axis lies in coarse interval [m[i], m[i] + 1] at relative position w[i]).
template <typename T> struct MgProl {
    int n0c = 0, n1c = 0;
    const int* mx = nullptr;
    const int* my = nullptr;
    const T* wx = nullptr;
    const T* wy = nullptr;
    const T* xc = nullptr;
};

template <typename T>
__device__ __forceinline__ T mg_prol_at(const MgProl<T>& pr, long long k, int i, int j) {
    const int I = pr.mx[i], J = pr.my[j];
    const int I1 = min(I + 1, pr.n0c - 1), J1 = min(J + 1, pr.n1c - 1);
    const T a = pr.wx[i], b = pr.wy[j];
    const T* c = pr.xc + k * pr.n0c * pr.n1c;
    const T lo = (T(1) - a) * c[(long long)J * pr.n0c + I] + a * c[(long long)J * pr.n0c + I1];
    const T hi = (T(1) - a) * c[(long long)J1 * pr.n0c + I] + a * c[(long long)J1 * pr.n0c + I1];
    return (T(1) - b) * lo + b * hi;
}

// Scalars of one PCG solve, resident on the device.  Every kernel of the
// iteration reads them; the last block of a reduction kernel writes them.
struct PcgState {
    double rz;       // (r, z) of the current iterate
    double rz_prev;  // (r, z) of the previous iterate
    double pap;      // (p, A p)
    double alpha;
    double rr;       // (r, r)
    double stop2;    // stop when rr <= stop2
    double bb;       // (b, b)
    double beta0;    // beta of the first iteration: 0, or (r,z)/(r,z)_prev when a run resumes
    int iters;
    int maxit;
    int done;    // 1: converged or stopped
    int status;  // 0 ok, 1 maxit reached, -6 (p,Ap) <= 0, -7 (r,z) <= 0, -8 non-finite
    unsigned int counter[4];  // block tickets for the last-block reductions
};

// dsh: a diagonal added to the operator, per node (the lumped heat-capacity matrix of
// the transient solver, C = diag(sum of the element capacities c_e over the elements
// touching the node), femT3d.cpp:233-242); nullptr for the steady operator.
template <typename T> struct Axes3 {
    int n0, n1, n2;
    const T* hx;  // element sizes along each axis (um)
    const T* hy;
    const T* hz;
    const T* dsh = nullptr;
};

template <typename T> struct Axes2 {
    int n0, n1;
    const T* hx;    // element sizes (um)
    const T* hy;
    const T* rmid;  // cylindrical: element mid radius, else nullptr
    const T* dsh = nullptr;
};

__device__ __forceinline__ bool dbit(const uint64_t* __restrict__ m, long long node) {
    return (m[node >> 6] >> (node & 63)) & 1ull;
}

// ---------------------------------------------------------------------------
// reductions: wave64 shuffle tree, LDS across waves, per-block partial, and the
// last block (ticket counter) sums the partials in fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Publish a block partial and take a ticket without an agent-scope release fence:
// on gfx950 a release fence writes back the XCD's whole L2, which made every block of a
// reduction kernel pay for the kernel's own streaming stores.  The partial is written
// with an agent-scope atomic exchange (performed at the device coherence point), the
// thread waits for it to complete (s_waitcnt vmcnt(0), also a compiler barrier), then
// increments the ticket; the last block reads the partials with agent-scope atomic loads.
__device__ __forceinline__ void publish_partial(double* p, double v) {
    __hip_atomic_exchange(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double read_partial(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ unsigned int take_ticket(unsigned int* counter) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    return __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-reduce NV values; returns true in thread 0 of the last block, which then
// holds the grid totals in tot[].
template <int NV>
__device__ bool grid_reduce(double (&v)[NV], double* __restrict__ partials, unsigned int* counter, double (&tot)[NV]) {
    __shared__ double red[NV][16];
    __shared__ bool last;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = (nthr + 63) >> 6;
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        double s = wave_sum(v[a]);
        if (lane == 0) red[a][wid] = s;
    }
    __syncthreads();
    const unsigned int nblk = gridDim.x * gridDim.y * gridDim.z;
    const unsigned int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (tid == 0) {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            double s = 0.;
            for (int w = 0; w < nw; ++w) s += red[a][w];
            publish_partial(partials + (size_t)a * nblk + bid, s);
        }
        last = (take_ticket(counter) == nblk - 1);
    }
    __syncthreads();
    if (!last) return false;
    // last block: sum partials in fixed order (deterministic)
#pragma unroll
    for (int a = 0; a < NV; ++a) {
        double s = 0.;
        for (unsigned int b = tid; b < nblk; b += nthr) s += read_partial(partials + (size_t)a * nblk + b);
        s = wave_sum(s);
        __syncthreads();
        if (lane == 0) red[a][wid] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.;
            for (int w = 0; w < nw; ++w) t += red[a][w];
            tot[a] = t;
        }
    }
    if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return tid == 0;
}

// ---------------------------------------------------------------------------
// 3D operator apply, z-marching.
//   MODE_OP : y = M A0 (M x) + (I-M) x        (the Dirichlet-eliminated operator)
//   MODE_RES: y = M (F - A0 x)                (residual of the full field x)
//   FUSEP   : x = z + beta p_old is formed on the fly from pin and stored into pout
//             (two buffers: halo nodes re-form p from pin while their owners write pout),
//             and the previous iteration's solution update xsol += alpha_prev p_old is
//             applied to owned nodes (p_old is read here anyway)
//   DOT     : (x, y) over owned free nodes -> PcgState.pap, alpha = rz / pap
// T is the storage and arithmetic type of vectors and element data.
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool FUSEP, bool DOT, int PFD = 1>
__global__ __launch_bounds__(1024) void k_apply3d(Axes3<T> g, int BY, int KC, int ytiles, int nb, int xcd,
                                                  const T* __restrict__ xin,
                                                  const T* __restrict__ zin, const T* __restrict__ pin,
                                                  T* __restrict__ pout, T* __restrict__ xsol, T* __restrict__ y,
                                                  const T* __restrict__ F, const uint64_t* __restrict__ dmask,
                                                  const T* __restrict__ kh, const T* __restrict__ kv,
                                                  PcgState* __restrict__ st, double* __restrict__ partials,
                                                  MgProl<T> pr) {
    extern __shared__ unsigned char smem_raw[];
    const int BX = g.n0;
    const int PX = BX + 1;  // padded row: column BX holds zeros
    T* xs = reinterpret_cast<T*>(smem_raw);       // [2][BY+1][PX]
    T* oc = xs + 2 * (BY + 1) * PX;               // [2][4][BY][PX], corner outputs at column tx+1
    if (FUSEP || DOT || MODE == MODE_RESG || MODE == MODE_RESP) {
        if (st->done) return;
    }
    // XCD-aware order: hardware block b runs on XCD b % 8; each XCD gets a contiguous
    // range of logical tiles (adjacent y tiles of a z chunk share halo rows in its L2).
    // Blocks past the last tile idle but still join the grid reduction.
    // xcd 2: logical tiles y-tile major (an XCD covers a range of in-plane rows over all z
    // chunks, the columns the multigrid sweeps of the same XCD handle).
    const int L = xcd ? (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const bool live = L < nb;
    const int tid = threadIdx.x;
    const int tx = tid % BX, ty = tid / BX;
    const bool active = live && ty < BY;
    const int nzc = (g.n2 + KC - 1) / KC;
    const int ytile = xcd == 2 ? L / nzc : L % ytiles, zchunk = xcd == 2 ? L % nzc : L / ytiles;
    const int jb = ytile * (BY - 1);
    const int j = jb - 1 + ty;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int kb = live ? zchunk * KC : n2;
    const int ke = min(kb + KC, n2);
    const bool nvalid = active && j >= 0 && j < n1;
    const bool owned = nvalid && ty >= 1;
    const bool evalid = active && tx < n0 - 1 && j >= 0 && j < n1 - 1;
    const bool xrow = active && ty == BY - 1 && (j + 1) < n1;  // also loads node row j+1
    const long long plane = (long long)n0 * n1;

    double beta = 0.;
    T alpha_prev = T(0);
    if (FUSEP) {
        beta = (st->iters == 0) ? st->beta0 : st->rz / st->rz_prev;
        alpha_prev = T(st->alpha);
    }

    // per-column geometric factors, folded with 1e-6 (W/m -> W/um) and 1/48
    T gxf = 0, gyf = 0, gzf = 0;
    if (evalid) {
        T hx = g.hx[tx], hy = g.hy[j];
        const T c = T(1e-6 / 48.);
        gxf = c * hy / hx;
        gyf = c * hx / hy;
        gzf = c * hx * hy;
    }
    // zero the padding column of both plane buffers and of the corner outputs
    for (int r = tid; r < 2 * (BY + 1); r += blockDim.x) xs[r * PX + BX] = T(0);
    for (int r = tid; r < 8 * BY; r += blockDim.x) oc[r * PX] = T(0);

    // Global loads of one step, issued one step ahead (software prefetch): node values
    // of plane k (own row and, for the last thread row, row j+1), the element data and
    // the height of layer k-1, and (residual modes) the load vector of plane k.  The
    // loads are branch-free -- lanes outside the grid read a clamped valid address and
    // ignore the value -- so that the compiler's wait counts before the next step's use
    // cover only this group and the prefetch stays in flight across the barrier.
    const uint32_t* dmask32 = reinterpret_cast<const uint32_t*>(dmask);
    const int jc = min(max(j, 0), n1 - 1), jxc = min(max(j + 1, 0), n1 - 1);
    const int ejc = min(max(j, 0), n1 - 2), etx = min(tx, n0 - 2);
    struct Raw {
        T z, p, zx, px, kxy, kz, x, f, hz, pc, pcx;
        uint32_t dw, dwx;
    };
    auto gload = [&](int k, Raw& R) {
        const int kc = min(k, n2 - 1);
        const long long node = (long long)kc * plane + (long long)jc * n0 + tx;
        if (FUSEP) {
            R.z = zin[node];
            R.p = pin[node];
            R.x = xsol[node];
        } else {
            R.z = xin[node];
        }
        if (MODE != MODE_OP) R.f = F[node];
        R.dw = dmask32[node >> 5];
        if constexpr (MODE == MODE_RESP) R.pc = mg_prol_at(pr, kc, tx, jc);
        if (ty == BY - 1) {
            const long long nx = (long long)kc * plane + (long long)jxc * n0 + tx;
            if (FUSEP) {
                R.zx = zin[nx];
                R.px = pin[nx];
            } else {
                R.zx = xin[nx];
            }
            R.dwx = dmask32[nx >> 5];
            if constexpr (MODE == MODE_RESP) R.pcx = mg_prol_at(pr, kc, tx, jxc);
        }
        const int ek = max(kc - 1, 0);
        const long long e = (long long)ek * (n0 - 1) * (n1 - 1) + (long long)ejc * (n0 - 1) + etx;
        R.kxy = kh[e];
        R.kz = kv[e];
        R.hz = g.hz[ek];
    };
    auto bit = [&](uint32_t w, int k, int jj) {
        const long long node = (long long)k * plane + (long long)jj * n0 + tx;
        return ((w >> (node & 31)) & 1u) != 0u;
    };
    // plane k values into xs[buf] (masked for MODE_OP), p_new stored for owned nodes
    auto lstore = [&](int k, int buf, const Raw& R, T& own, bool& own_dir, T& own_f, bool store_p) {
        T v = T(0);
        bool d = false;
        if (nvalid) {
            v = FUSEP ? R.z + T(beta) * R.p : R.z;
            d = bit(R.dw, k, j);
            if constexpr (MODE == MODE_RESP) {
                if (!d) v += R.pc;
            }
            if (FUSEP && store_p) {
                const long long node = (long long)k * plane + (long long)j * n0 + tx;
                pout[node] = v;
                xsol[node] = R.x + alpha_prev * R.p;
            }
        }
        own = v;
        own_dir = d;
        own_f = R.f;
        if (active) xs[(buf * (BY + 1) + ty) * PX + tx] = (MODE == MODE_OP && d) ? T(0) : v;
        if (active && ty == BY - 1) {
            T w = T(0);
            if (xrow) {
                w = FUSEP ? R.zx + T(beta) * R.px : R.zx;
                if (MODE == MODE_OP && bit(R.dwx, k, j + 1)) w = T(0);
                if constexpr (MODE == MODE_RESP) {
                    if (!bit(R.dwx, k, j + 1)) w += R.pcx;
                }
            }
            xs[(buf * (BY + 1) + BY) * PX + tx] = w;
        }
    };
    // in-plane sum/difference transform of the element column's 4 corners
    auto plane_tf = [&](int buf, T (&t)[4]) {
        const T* r0 = xs + (buf * (BY + 1) + ty) * PX;
        const T* r1 = r0 + PX;
        T v00 = r0[tx], v10 = r0[tx + 1], v01 = r1[tx], v11 = r1[tx + 1];
        T s0 = v00 + v10, d0 = v10 - v00, s1 = v01 + v11, d1 = v11 - v01;
        t[0] = s0 + s1;  // sigma sigma
        t[1] = d0 + d1;  // delta sigma   (x delta)
        t[2] = s1 - s0;  // sigma delta   (y delta)
        t[3] = d1 - d0;  // delta delta
    };

    const int P0 = !live ? n2 : (kb > 0) ? kb - 1 : 0;
    int buf = 0;
    T own_cur, own_next = T(0), f_cur = T(0), f_next = T(0);
    bool dir_cur, dir_next = false;
    // PFD planes in flight: R1 = plane P + 1, R2 = plane P + 2 (PFD == 2)
    Raw R1, R2;
    R1.z = R1.p = R1.zx = R1.px = R1.kxy = R1.kz = R1.x = R1.f = R1.hz = R1.pc = R1.pcx = T(0);
    R1.dw = R1.dwx = 0u;
    R2 = R1;
    __syncthreads();
    if (live) {
        gload(P0, R1);
        lstore(P0, 0, R1, own_cur, dir_cur, f_cur, FUSEP && P0 >= kb && owned);
        gload(P0 + 1, R1);
        if (PFD == 2) gload(P0 + 2, R2);
    }
    __syncthreads();
    T xt_cur[4], xt_next[4] = {T(0), T(0), T(0), T(0)};
    if (active) plane_tf(0, xt_cur);
    T acc[4] = {T(0), T(0), T(0), T(0)};
    double dot = 0.;

    // One barrier per plane: the corner contributions of plane P go to one of two
    // corner buffers (alternating by step) after the barrier of step P and are
    // gathered after the barrier of step P + 1.
    T own_prev = T(0), f_prev = T(0);
    bool dir_prev = false;
    int obuf = 0;
    auto gather = [&](int Pg, int ob, T own_g, bool dir_g, T f_g) {
        const T* o = oc + ob * 4 * BY * PX;
        T s = o[(0 * BY + ty) * PX + tx + 1] + o[(1 * BY + ty) * PX + tx] + o[(2 * BY + ty - 1) * PX + tx + 1] +
              o[(3 * BY + ty - 1) * PX + tx];
        long long node = (long long)Pg * plane + (long long)j * n0 + tx;
        if (g.dsh) s += g.dsh[node] * own_g;
        T out;
        if (MODE == MODE_OP) {
            out = dir_g ? own_g : s;
        } else {
            out = dir_g ? T(0) : (f_g - s);
        }
        y[node] = out;
        if (DOT && !dir_g) dot += (double)own_g * (double)out;
    };
    for (int P = P0; P < ke; ++P) {
        const bool has_next = (P + 1) < n2;
        T bot[4] = {T(0), T(0), T(0), T(0)}, top[4] = {T(0), T(0), T(0), T(0)};
        T a = T(0), b = T(0), c = T(0);
        if (has_next) {
            lstore(P + 1, buf ^ 1, R1, own_next, dir_next, f_next, FUSEP && (P + 1) >= kb && (P + 1) < ke && owned);
            if (evalid) {
                const T hz = R1.hz;
                a = R1.kxy * (gxf * hz);
                b = R1.kxy * (gyf * hz);
                c = R1.kz * (gzf / hz);
            }
        }
        // prefetch (clamped planes past the chunk's top halo are read and unused)
        if (PFD == 2) {
            R1 = R2;
            gload(P + 3, R2);
        } else {
            gload(P + 2, R1);
        }
        __syncthreads();
        // outputs of plane P - 1, whose corner contributions were written in the last step
        if (P - 1 >= kb && owned) gather(P - 1, obuf ^ 1, own_prev, dir_prev, f_prev);
        if (has_next && active) {
            plane_tf(buf ^ 1, xt_next);
            // z stage: w[pq sigma] = T + B, w[pq delta] = T - B, scaled by the eigenvalues
            const T ab = a + b, ac = a + c, bc = b + c, abc3 = (ab + c) * T(1. / 3.);
            // components: 0 = ss, 1 = ds, 2 = sd, 3 = dd (x,y); z sigma / delta
            T ws1 = (xt_next[1] + xt_cur[1]) * (T(3) * a);   // d s s
            T ws2 = (xt_next[2] + xt_cur[2]) * (T(3) * b);   // s d s
            T ws3 = (xt_next[3] + xt_cur[3]) * ab;           // d d s
            T wd0 = (xt_next[0] - xt_cur[0]) * (T(3) * c);   // s s d
            T wd1 = (xt_next[1] - xt_cur[1]) * ac;           // d s d
            T wd2 = (xt_next[2] - xt_cur[2]) * bc;           // s d d
            T wd3 = (xt_next[3] - xt_cur[3]) * abc3;         // d d d
            bot[0] = -wd0;
            top[0] = wd0;
            bot[1] = ws1 - wd1;
            top[1] = ws1 + wd1;
            bot[2] = ws2 - wd2;
            top[2] = ws2 + wd2;
            bot[3] = ws3 - wd3;
            top[3] = ws3 + wd3;
        }
        if (P >= kb && active) {
            // plane P is complete: inverse in-plane transform -> 4 corner outputs
            T* o = oc + obuf * 4 * BY * PX;
            T t0 = acc[0] + bot[0], t1 = acc[1] + bot[1], t2 = acc[2] + bot[2], t3 = acc[3] + bot[3];
            T as = t0 - t2, ad = t1 - t3;  // row j   (y sigma - y delta)
            T bs = t0 + t2, bd = t1 + t3;  // row j+1
            o[(0 * BY + ty) * PX + tx + 1] = as - ad;  // corner (0,0)
            o[(1 * BY + ty) * PX + tx + 1] = as + ad;  // corner (1,0)
            o[(2 * BY + ty) * PX + tx + 1] = bs - bd;  // corner (0,1)
            o[(3 * BY + ty) * PX + tx + 1] = bs + bd;  // corner (1,1)
        }
        for (int q = 0; q < 4; ++q) {
            acc[q] = top[q];
            xt_cur[q] = xt_next[q];
        }
        own_prev = own_cur;
        dir_prev = dir_cur;
        f_prev = f_cur;
        own_cur = own_next;
        dir_cur = dir_next;
        f_cur = f_next;
        buf ^= 1;
        obuf ^= 1;
    }
    // drain: outputs of the last plane
    __syncthreads();
    if (ke - 1 >= kb && owned) gather(ke - 1, obuf ^ 1, own_prev, dir_prev, f_prev);
    if (DOT) {
        double v[1] = {dot}, tot[1];
        if (grid_reduce<1>(v, partials, &st->counter[0], tot)) {
            st->pap = tot[0];
            if (!(tot[0] > 0.) || !isfinite(tot[0])) {
                st->status = isfinite(tot[0]) ? -6 : -8;
                st->done = 1;
                st->alpha = 0.;
            } else {
                st->alpha = st->rz / tot[0];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 3D operator apply, row blocks: one block per node row (y index j) and chunk of
// KR planes, thread = x index, marching along z.  Modes as k_apply3d.  Each
// thread forms the element columns (i, j-1) and (i, j) (rows j-1, j, j+1 of
// nodes; the neighbour rows are other blocks' own rows, read through L2), so a
// block needs no y halo exchange: per plane it shares only the x+1 node values
// and the x+1 corner contributions through LDS (two barriers of a 1-4 wave
// block, which many co-resident blocks overlap).  XCD-aware block order:
// hardware block b runs on XCD b % 8, and each XCD gets a contiguous range of
// logical blocks (adjacent rows), so neighbour rows hit in that XCD's L2.
// Blocks have round64(n0) <= 512 threads.  The grid is a multiple of 8; blocks
// past the last row idle but still join
// the grid reduction.
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool FUSEP, bool DOT>
__global__ __launch_bounds__(512) void k_apply3r(Axes3<T> g, int KR, int nb, const T* __restrict__ xin,
                                                  const T* __restrict__ zin, const T* __restrict__ pin,
                                                  T* __restrict__ pout, T* __restrict__ xsol, T* __restrict__ y,
                                                  const T* __restrict__ F, const uint64_t* __restrict__ dmask,
                                                  const T* __restrict__ kh, const T* __restrict__ kv,
                                                  PcgState* __restrict__ st, double* __restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int PX = n0 + 1;
    T* xs = reinterpret_cast<T*>(smem_raw);  // [3][PX]: rows j-1, j, j+1 of the incoming plane
    T* oc = xs + 3 * PX;                     // [PX]: corner contributions for node i + 1
    if (FUSEP || DOT || MODE == MODE_RESG) {
        if (st->done) return;
    }
    const int q8 = gridDim.x >> 3;
    const int L = (int)(blockIdx.x & 7) * q8 + (int)(blockIdx.x >> 3);
    const bool live = L < nb;
    double dot = 0.;
    if (live) {
        const int j = L % n1, zc = L / n1;
        const int kb = zc * KR, ke = min(kb + KR, n2);
        const int i = threadIdx.x;
        const bool ival = i < n0;
        const bool ev = i < n0 - 1;
        const bool hasA = ev && j >= 1, hasB = ev && j < n1 - 1;
        const long long plane = (long long)n0 * n1;
        const long long eplane = (long long)(n0 - 1) * (n1 - 1);

        double beta = 0.;
        T alpha_prev = T(0);
        if (FUSEP) {
            beta = (st->iters == 0) ? st->beta0 : st->rz / st->rz_prev;
            alpha_prev = T(st->alpha);
        }
        const T c48 = T(1e-6 / 48.);
        T gxA = 0, gyA = 0, gzA = 0, gxB = 0, gyB = 0, gzB = 0;
        if (hasA) {
            const T hx = g.hx[i], hy = g.hy[j - 1];
            gxA = c48 * hy / hx;
            gyA = c48 * hx / hy;
            gzA = c48 * hx * hy;
        }
        if (hasB) {
            const T hx = g.hx[i], hy = g.hy[j];
            gxB = c48 * hy / hx;
            gyB = c48 * hx / hy;
            gzB = c48 * hx * hy;
        }
        for (int r = threadIdx.x; r < 3; r += blockDim.x) xs[r * PX + n0] = T(0);
        if (threadIdx.x == 0) oc[0] = T(0);

        struct Raw {
            T v[3], p[3], x, axy, az, bxy, bz;
            unsigned d;
        };
        auto gload = [&](int k, Raw& R) {
            R.d = 0;
            if (ival) {
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const int row = j - 1 + r;
                    if (row >= 0 && row < n1) {
                        const long long node = (long long)k * plane + (long long)row * n0 + i;
                        if (FUSEP) {
                            R.v[r] = zin[node];
                            R.p[r] = pin[node];
                        } else {
                            R.v[r] = xin[node];
                        }
                        if (dbit(dmask, node)) R.d |= 1u << r;
                    }
                }
                if (FUSEP && k >= kb && k < ke) R.x = xsol[(long long)k * plane + (long long)j * n0 + i];
            }
            if (k >= 1) {
                const long long e0 = (long long)(k - 1) * eplane + i;
                if (hasA) {
                    R.axy = kh[e0 + (long long)(j - 1) * (n0 - 1)];
                    R.az = kv[e0 + (long long)(j - 1) * (n0 - 1)];
                }
                if (hasB) {
                    R.bxy = kh[e0 + (long long)j * (n0 - 1)];
                    R.bz = kv[e0 + (long long)j * (n0 - 1)];
                }
            }
        };
        // node values of plane k (rows j-1, j, j+1) into xs; own value and Dirichlet flag out
        auto lstore = [&](int k, const Raw& R, T& own, bool& own_dir, bool store) {
            T m[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int row = j - 1 + r;
                T v = T(0);
                if (ival && row >= 0 && row < n1) v = FUSEP ? R.v[r] + T(beta) * R.p[r] : R.v[r];
                if (r == 1) {
                    own = v;
                    own_dir = (R.d >> 1) & 1u;
                    if (FUSEP && store && ival) {
                        const long long node = (long long)k * plane + (long long)j * n0 + i;
                        pout[node] = v;
                        xsol[node] = R.x + alpha_prev * R.p[1];
                    }
                }
                m[r] = (MODE == MODE_OP && ((R.d >> r) & 1u)) ? T(0) : v;
            }
            if (ival) {
#pragma unroll
                for (int r = 0; r < 3; ++r) xs[r * PX + i] = m[r];
            }
        };
        // in-plane sum/difference transforms of element columns A = (i, j-1), B = (i, j)
        auto plane_tf = [&](T (&tA)[4], T (&tB)[4]) {
            T v[3], w[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                v[r] = ival ? xs[r * PX + i] : T(0);
                w[r] = ival ? xs[r * PX + i + 1] : T(0);
            }
            // A: lower row j-1 (r = 0), upper row j (r = 1); B: rows j, j+1
            T s0 = v[0] + w[0], d0 = w[0] - v[0], s1 = v[1] + w[1], d1 = w[1] - v[1];
            T s2 = v[2] + w[2], d2 = w[2] - v[2];
            tA[0] = s0 + s1;
            tA[1] = d0 + d1;
            tA[2] = s1 - s0;
            tA[3] = d1 - d0;
            tB[0] = s1 + s2;
            tB[1] = d1 + d2;
            tB[2] = s2 - s1;
            tB[3] = d2 - d1;
        };
        // z stage of one element column (eigenvalues of the Q1 box, a = kx'/48 ...)
        auto zstage = [&](const T (&tc)[4], const T (&tn)[4], T a, T b, T c, T (&bot)[4], T (&top)[4]) {
            const T ab = a + b, ac = a + c, bc = b + c, abc3 = (ab + c) * T(1. / 3.);
            const T ws1 = (tn[1] + tc[1]) * (T(3) * a);
            const T ws2 = (tn[2] + tc[2]) * (T(3) * b);
            const T ws3 = (tn[3] + tc[3]) * ab;
            const T wd0 = (tn[0] - tc[0]) * (T(3) * c);
            const T wd1 = (tn[1] - tc[1]) * ac;
            const T wd2 = (tn[2] - tc[2]) * bc;
            const T wd3 = (tn[3] - tc[3]) * abc3;
            bot[0] = -wd0;
            top[0] = wd0;
            bot[1] = ws1 - wd1;
            top[1] = ws1 + wd1;
            bot[2] = ws2 - wd2;
            top[2] = ws2 + wd2;
            bot[3] = ws3 - wd3;
            top[3] = ws3 + wd3;
        };

        const int P0 = kb > 0 ? kb - 1 : 0;
        T own_cur, own_next = T(0);
        bool dir_cur, dir_next = false;
        Raw R;
#pragma unroll
        for (int r = 0; r < 3; ++r) R.v[r] = R.p[r] = T(0);
        R.x = R.axy = R.az = R.bxy = R.bz = T(0);
        R.d = 0;
        gload(P0, R);
        lstore(P0, R, own_cur, dir_cur, FUSEP && P0 >= kb);
        if (P0 + 1 < n2) gload(P0 + 1, R);
        __syncthreads();
        T tAc[4], tBc[4], tAn[4] = {T(0), T(0), T(0), T(0)}, tBn[4] = {T(0), T(0), T(0), T(0)};
        plane_tf(tAc, tBc);
        T accA[4] = {T(0), T(0), T(0), T(0)}, accB[4] = {T(0), T(0), T(0), T(0)};
        __syncthreads();
        for (int P = P0; P < ke; ++P) {
            const bool has_next = (P + 1) < n2;
            T botA[4] = {T(0), T(0), T(0), T(0)}, topA[4] = {T(0), T(0), T(0), T(0)};
            T botB[4] = {T(0), T(0), T(0), T(0)}, topB[4] = {T(0), T(0), T(0), T(0)};
            if (has_next) {
                lstore(P + 1, R, own_next, dir_next, FUSEP && (P + 1) >= kb && (P + 1) < ke);
                const T hz = g.hz[P];
                const T aA = R.axy * (gxA * hz), bA = R.axy * (gyA * hz), cA = R.az * (gzA / hz);
                const T aB = R.bxy * (gxB * hz), bB = R.bxy * (gyB * hz), cB = R.bz * (gzB / hz);
                if (P + 2 < n2 && P + 2 <= ke) gload(P + 2, R);  // prefetch the next step
                __syncthreads();
                plane_tf(tAn, tBn);
                zstage(tAc, tAn, aA, bA, cA, botA, topA);
                zstage(tBc, tBn, aB, bB, cB, botB, topB);
            } else {
                __syncthreads();
            }
            T own_c = T(0);
            if (P >= kb) {
                const T a0 = accA[0] + botA[0], a1 = accA[1] + botA[1], a2 = accA[2] + botA[2],
                        a3 = accA[3] + botA[3];
                const T b0 = accB[0] + botB[0], b1 = accB[1] + botB[1], b2 = accB[2] + botB[2],
                        b3 = accB[3] + botB[3];
                // A: corners on its upper row (j); B: corners on its lower row (j)
                const T As = a0 + a2, Ad = a1 + a3;
                const T Bs = b0 - b2, Bd = b1 - b3;
                own_c = (As - Ad) + (Bs - Bd);
                if (ival) oc[i + 1] = (As + Ad) + (Bs + Bd);
            }
            __syncthreads();
            if (P >= kb && ival) {
                const long long node = (long long)P * plane + (long long)j * n0 + i;
                const T s = own_c + oc[i] + (g.dsh ? g.dsh[node] * own_cur : T(0));
                T out;
                if (MODE == MODE_OP) out = dir_cur ? own_cur : s;
                else out = dir_cur ? T(0) : (F[node] - s);
                y[node] = out;
                if (DOT && !dir_cur) dot += (double)own_cur * (double)out;
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                accA[a] = topA[a];
                accB[a] = topB[a];
                tAc[a] = tAn[a];
                tBc[a] = tBn[a];
            }
            own_cur = own_next;
            dir_cur = dir_next;
        }
    }
    if (DOT) {
        double v[1] = {dot}, tot[1];
        if (grid_reduce<1>(v, partials, &st->counter[0], tot)) {
            st->pap = tot[0];
            if (!(tot[0] > 0.) || !isfinite(tot[0])) {
                st->status = isfinite(tot[0]) ? -6 : -8;
                st->done = 1;
                st->alpha = 0.;
            } else {
                st->alpha = st->rz / tot[0];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 2D operator apply (Cartesian or cylindrical), marching along axis 1.
// One block = the whole row (BX = n0 threads); thread tx owns element column tx
// and node column tx.  Element matrix therm2d.cpp:216-228, x r_mid for
// cylindrical geometry (therm2d.cpp:333,410-420; electr2d.cpp:206-231).
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool FUSEP, bool DOT>
__global__ __launch_bounds__(1024) void k_apply2d(Axes2<T> g, int KC, const T* __restrict__ xin,
                                                  const T* __restrict__ zin, const T* __restrict__ pin,
                                                  T* __restrict__ pout, T* __restrict__ xsol, T* __restrict__ y,
                                                  const T* __restrict__ F, const uint64_t* __restrict__ dmask,
                                                  const T* __restrict__ kh, const T* __restrict__ kv,
                                                  PcgState* __restrict__ st, double* __restrict__ partials) {
    extern __shared__ unsigned char smem_raw[];
    const int n0 = g.n0, n1 = g.n1;
    const int PX = n0 + 1;
    T* xs = reinterpret_cast<T*>(smem_raw);  // [2][PX]
    T* oc = xs + 2 * PX;                     // [2][PX], corner outputs at column tx+1
    if (FUSEP || DOT || MODE == MODE_RESG) {
        if (st->done) return;
    }
    const int tx = threadIdx.x;
    const bool nvalid = tx < n0;
    const bool evalid = tx < n0 - 1;
    const int kb = blockIdx.x * KC;
    const int ke = min(kb + KC, n1);
    double beta = 0.;
    T alpha_prev = T(0);
    if (FUSEP) {
        beta = (st->iters == 0) ? st->beta0 : st->rz / st->rz_prev;
        alpha_prev = T(st->alpha);
    }

    T gxf = 0, gyf = 0;
    if (evalid) {
        T w = g.hx[tx];
        T r = g.rmid ? g.rmid[tx] : T(1);
        gxf = r / (T(12) * w);  // kx' / 12 = kx h r / (12 w)
        gyf = r * w / T(12);    // ky' / 12 = ky w r / (12 h)
    }
    if (threadIdx.x < 2) {
        xs[threadIdx.x * PX + n0] = T(0);
        oc[threadIdx.x * PX] = T(0);
    }
    auto load_row = [&](int k, int buf, T& own, bool& d, bool store_p) {
        T v = T(0);
        d = false;
        if (nvalid) {
            long long node = (long long)k * n0 + tx;
            if (FUSEP) {
                const T po = pin[node];
                v = zin[node] + T(beta) * po;
                if (store_p) {
                    pout[node] = v;
                    xsol[node] += alpha_prev * po;
                }
            } else {
                v = xin[node];
            }
            d = dbit(dmask, node);
            xs[buf * PX + tx] = (MODE == MODE_OP && d) ? T(0) : v;
        }
        own = v;
    };
    const int P0 = (kb > 0) ? kb - 1 : 0;
    int buf = 0;
    T own_cur, own_next = T(0);
    bool dir_cur, dir_next = false;
    __syncthreads();
    load_row(P0, 0, own_cur, dir_cur, FUSEP && P0 >= kb);
    __syncthreads();
    T s_cur = T(0), d_cur = T(0);
    if (nvalid) {
        T v0 = xs[tx], v1 = xs[tx + 1];
        s_cur = v0 + v1;
        d_cur = v1 - v0;
    }
    T acc0 = T(0), acc1 = T(0);
    double dot = 0.;
    for (int P = P0; P < ke; ++P) {
        const bool has_next = (P + 1) < n1;
        T bot0 = T(0), bot1 = T(0), top0 = T(0), top1 = T(0);
        T s_next = T(0), d_next = T(0);
        if (has_next) {
            load_row(P + 1, buf ^ 1, own_next, dir_next, FUSEP && (P + 1) < ke);
            T a = T(0), b = T(0);
            if (evalid) {
                long long e = (long long)P * (n0 - 1) + tx;
                T h = g.hy[P];
                a = kh[e] * (gxf * h);
                b = kv[e] * (gyf / h);
            }
            __syncthreads();
            if (nvalid) {
                T v0 = xs[(buf ^ 1) * PX + tx], v1 = xs[(buf ^ 1) * PX + tx + 1];
                s_next = v0 + v1;
                d_next = v1 - v0;
                // w[x p][z q]: (s,s) eigenvalue 0, (d,s) 3a, (s,d) 3b, (d,d) a+b
                T wds = (d_next + d_cur) * (T(3) * a);
                T wsd = (s_next - s_cur) * (T(3) * b);
                T wdd = (d_next - d_cur) * (a + b);
                bot0 = -wsd;
                top0 = wsd;
                bot1 = wds - wdd;
                top1 = wds + wdd;
            }
        }
        if (P >= kb) {
            if (nvalid) {
                T t0 = acc0 + bot0, t1 = acc1 + bot1;
                oc[0 * PX + tx + 1] = t0 - t1;  // corner at node tx
                oc[1 * PX + tx + 1] = t0 + t1;  // corner at node tx+1
            }
            __syncthreads();
            if (nvalid) {
                long long node = (long long)P * n0 + tx;
                T s = oc[0 * PX + tx + 1] + oc[1 * PX + tx];
                if (g.dsh) s += g.dsh[node] * own_cur;
                T out;
                if (MODE == MODE_OP) out = dir_cur ? own_cur : s;
                else out = dir_cur ? T(0) : (F[node] - s);
                y[node] = out;
                if (DOT && !dir_cur) dot += (double)own_cur * (double)out;
            }
        }
        acc0 = top0;
        acc1 = top1;
        s_cur = s_next;
        d_cur = d_next;
        own_cur = own_next;
        dir_cur = dir_next;
        buf ^= 1;
    }
    if (DOT) {
        double v[1] = {dot}, tot[1];
        if (grid_reduce<1>(v, partials, &st->counter[0], tot)) {
            st->pap = tot[0];
            if (!(tot[0] > 0.) || !isfinite(tot[0])) {
                st->status = isfinite(tot[0]) ? -6 : -8;
                st->done = 1;
                st->alpha = 0.;
            } else {
                st->alpha = st->rz / tot[0];
            }
        }
    }
}

// end of an iteration: new (r,z), (r,r) -> scalars, stopping decisions
__device__ inline void finish_update(PcgState* st, const double (&tot)[2]) {
    st->rz_prev = st->rz;
    st->rz = tot[0];
    st->rr = tot[1];
    st->iters += 1;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        st->status = -8;
        st->done = 1;
    } else if (tot[1] <= st->stop2) {
        st->done = 1;
    } else if (!(tot[0] > 0.)) {
        st->status = -7;
        st->done = 1;
    } else if (st->iters >= st->maxit) {
        st->status = 1;
        st->done = 1;
    }
}

// start of a CG run: scalars from (r,z), (r,r)
// Start of a CG run.  beta0 > 0 on entry marks a resumed run (the residual was replaced
// and the previous direction kept): it then holds the previous (r, z) in the current
// scaling and becomes the first beta.
__device__ inline void finish_start(PcgState* st, const double (&tot)[2]) {
    st->rz = tot[0];
    st->rz_prev = tot[0];
    st->beta0 = st->beta0 > 0. ? tot[0] / st->beta0 : 0.;
    st->rr = tot[1];
    st->iters = 0;
    st->status = 0;
    st->done = (tot[1] <= st->stop2) ? 1 : 0;
    if (!isfinite(tot[1])) {
        st->status = -8;
        st->done = 1;
    }
}

// ---------------------------------------------------------------------------
// PCG vector step: r -= alpha q; z = Dinv r; (r,z), (r,r).  (x += alpha p is applied
// by the next fused apply, which reads p anyway, or by k_x_fixup after the last one.)
// Grid-stride over all nodes; vectors of type T, reductions in FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(long long n, T* __restrict__ r, const T* __restrict__ q,
                                                   const T* __restrict__ dinv, T* __restrict__ z,
                                                   PcgState* __restrict__ st, double* __restrict__ partials) {
    if (st->done) return;
    const T alpha = T(st->alpha);
    double rz = 0., rr = 0.;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        T qi = q[i];
        T ri = r[i] - alpha * qi;
        T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += (double)ri * (double)ri;
    }
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2>(v, partials, &st->counter[1], tot)) finish_update(st, tot);
}

// ---------------------------------------------------------------------------
// Boundary matrix: the off-diagonal part of a sparse symmetric addition to the operator
// (convection faces), rows in CSR over the nodes it touches; its diagonal travels with the
// operator's diagonal shift.  MODE_OP: y += M Kb M x; MODE_RES: y -= M Kb x.  Skipped
// after the stop when st is given (inside a CG run).
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_bnd_apply(long long nr, const long long* __restrict__ rows,
                                                   const int* __restrict__ ptr, const long long* __restrict__ cols,
                                                   const T* __restrict__ vals, const uint64_t* __restrict__ dmask,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   const PcgState* __restrict__ st) {
    if (st && st->done) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nr; i += (long long)gridDim.x * blockDim.x) {
        const long long n = rows[i];
        if (dbit(dmask, n)) continue;
        T s = T(0);
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const long long c = cols[k];
            if (MODE == MODE_OP && dbit(dmask, c)) continue;
            s += vals[k] * x[c];
        }
        if (MODE == MODE_OP)
            y[n] += s;
        else
            y[n] -= s;
    }
}

// Unfused CG direction step (used with a boundary matrix): p = z + beta p_old and the
// previous iteration's x += alpha_prev p_old, the work the fused apply does on the fly
template <typename T>
__global__ __launch_bounds__(256) void k_cg_pstep(long long n, const T* __restrict__ z, const T* __restrict__ pin,
                                                  T* __restrict__ pout, T* __restrict__ x,
                                                  const PcgState* __restrict__ st) {
    if (st->done) return;
    const T beta = T((st->iters == 0) ? st->beta0 : st->rz / st->rz_prev);
    const T ap = T(st->alpha);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const T po = pin[i];
        pout[i] = z[i] + beta * po;
        x[i] += ap * po;
    }
}

// (p, q) over the free nodes -> pap, alpha = rz / pap (the fused apply's DOT step)
template <typename T>
__global__ __launch_bounds__(256) void k_cg_dot(long long n, const T* __restrict__ p, const T* __restrict__ q,
                                                const uint64_t* __restrict__ dmask, PcgState* __restrict__ st,
                                                double* __restrict__ partials) {
    if (st->done) return;
    double d = 0.;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!dbit(dmask, i)) d += (double)p[i] * (double)q[i];
    double v[1] = {d}, tot[1];
    if (grid_reduce<1>(v, partials, &st->counter[0], tot)) {
        st->pap = tot[0];
        if (!(tot[0] > 0.) || !isfinite(tot[0])) {
            st->status = isfinite(tot[0]) ? -6 : -8;
            st->done = 1;
            st->alpha = 0.;
        } else {
            st->alpha = st->rz / tot[0];
        }
    }
}

// node diagonal shift = capacity + boundary diagonal (either may be absent)
__global__ void k_sum_dsh(long long n, const double* __restrict__ a, const double* __restrict__ b,
                          double* __restrict__ d, float* __restrict__ d32) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double v = (a ? a[i] : 0.) + (b ? b[i] : 0.);
        d[i] = v;
        d32[i] = (float)v;
    }
}

// last solution update after the iteration stopped: x += alpha p_last, p_last in
// pb[iters & 1] (the buffer the last fused apply wrote)
template <typename T>
__global__ void k_x_fixup(long long n, T* __restrict__ x, const T* __restrict__ p0, const T* __restrict__ p1,
                          const PcgState* __restrict__ st) {
    const T alpha = T(st->alpha);
    const T* p = (st->iters & 1) ? p1 : p0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] += alpha * p[i];
}

// r = given residual; z = Dinv r; (r,z), (r,r) -> state (start of a CG run)
template <typename T>
__global__ __launch_bounds__(256) void k_cg_start(long long n, const T* __restrict__ r, const T* __restrict__ dinv,
                                                  T* __restrict__ z, PcgState* __restrict__ st,
                                                  double* __restrict__ partials) {
    double rz = 0., rr = 0.;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        T ri = r[i];
        T zi = dinv[i] * ri;
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += (double)ri * (double)ri;
    }
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2>(v, partials, &st->counter[1], tot)) finish_start(st, tot);
}

// FP64 PCG with an FP32 preconditioner, before it: r -= alpha q (FP64, skipped when
// upd is false) and the FP32 copy r32 of the residual.  Exits after the stop.
__global__ __launch_bounds__(256) void k_cg_r_to_f32(long long n, bool upd, double* __restrict__ r,
                                                     const double* __restrict__ q, float* __restrict__ r32,
                                                     const PcgState* __restrict__ st) {
    if (st->done) return;
    const double alpha = upd ? st->alpha : 0.;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double ri = r[i];
        if (upd) {
            ri -= alpha * q[i];
            r[i] = ri;
        }
        r32[i] = (float)ri;
    }
}

// ... and after it: z = z32 in FP64, (r, z), (r, r); fin 1 ends the iteration, 2 starts a run
__global__ __launch_bounds__(256) void k_cg_z_from_f32(long long n, int fin, const double* __restrict__ r,
                                                       const float* __restrict__ z32, double* __restrict__ z,
                                                       PcgState* __restrict__ st, double* __restrict__ partials) {
    if (st->done) return;
    double rz = 0., rr = 0.;
    // four nodes a grid apart per step, their loads issued together (one at a time, the
    // loop left this kernel at 3 TB/s; more blocks instead cost more in the grid
    // reduction's ticket atomics than they gained)
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        double ri[4], zi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ri[u] = r[i + u * stride];
            zi[u] = (double)z32[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            z[i + u * stride] = zi[u];
            rz += ri[u] * zi[u];
            rr += ri[u] * ri[u];
        }
    }
    for (; i < n; i += stride) {
        const double ri = r[i], zi = (double)z32[i];
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2>(v, partials, &st->counter[1], tot)) {
        if (fin == 2) finish_start(st, tot);
        else finish_update(st, tot);
    }
}

// sum of squares of a vector -> st->bb
template <typename T>
__global__ __launch_bounds__(256) void k_norm2(long long n, const T* __restrict__ v_, PcgState* __restrict__ st,
                                               double* __restrict__ partials, int slot) {
    double s = 0.;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double v = (double)v_[i];
        s += v * v;
    }
    double v[1] = {s}, tot[1];
    if (grid_reduce<1>(v, partials, &st->counter[2], tot)) {
        if (slot == 0) st->bb = tot[0];
        else st->rr = tot[0];
    }
}

// ---------------------------------------------------------------------------
// Line (block-Jacobi) preconditioner along the last axis (the layer-stacking
// axis): M = the tridiagonal part of A coupling each node to its two neighbours
// in the same column.  Thin layers make those couplings dominate the operator,
// so M captures most of it.  Factors of the Thomas algorithm are precomputed:
//   w_k = 1 / (b_k - a_k c'_{k-1}),  c'_k = c_k w_k,  a_k = c_{k-1}
// and z = M^-1 r is  d_k = (r_k - a_k d_{k-1}) w_k,  z_k = d_k - c'_k z_{k+1}.
// Each line block is a principal submatrix of the SPD matrix, so all pivots
// are positive.
// ---------------------------------------------------------------------------
// Per-plane diagonal b_k and coupling c_k (row k to row k+1) of the line matrix of
// one column, with Dirichlet rows reduced to the identity and their couplings
// dropped (the symmetric elimination of A).
// 3D: lines along axis 2; b = sum_e (kx'+ky'+kz')/9, c = sum (kx'+ky'-2kz')/18
// over the 4 elements holding the z edge (therm3d.cpp:227-237).
struct LineBC3 {
    Axes3<double> g;
    const double* kh;
    const double* kv;
    const uint64_t* dmask;
    int i, j;
    long long plane;
    __device__ void operator()(int k, long long node, double& b, double& c) const {
        b = 0.;
        c = 0.;
        for (int cc = 0; cc < 8; ++cc) {
            int ei = i - 1 + (cc & 1), ej = j - 1 + ((cc >> 1) & 1), ek = k - 1 + ((cc >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= g.n0 - 1 || ej >= g.n1 - 1 || ek >= g.n2 - 1) continue;
            long long e = ((long long)ek * (g.n1 - 1) + ej) * (g.n0 - 1) + ei;
            double dx = g.hx[ei], dy = g.hy[ej], dz = g.hz[ek];
            double kxy = kh[e] * 1e-6, kz = kv[e] * 1e-6;
            double kx1 = kxy * dy * dz / dx, ky1 = kxy * dx * dz / dy, kz1 = kz * dx * dy / dz;
            b += (kx1 + ky1 + kz1) / 9.;
            if (ek == k) c += (kx1 + ky1 - 2. * kz1) / 18.;  // z edge (k, k+1) of this element
        }
        if (g.dsh) b += g.dsh[node];
        const bool dk = dbit(dmask, node);
        if (dk) b = 1.;
        if (dk || k == g.n2 - 1 || dbit(dmask, node + plane)) c = 0.;
    }
};

// 2D: lines along axis 1; b = sum (kx'+ky')/3, c = sum (kx'-2ky')/6 (x r_mid),
// therm2d.cpp:225-228
struct LineBC2 {
    Axes2<double> g;
    const double* kh;
    const double* kv;
    const uint64_t* dmask;
    int i;
    __device__ void operator()(int k, long long node, double& b, double& c) const {
        b = 0.;
        c = 0.;
        for (int cc = 0; cc < 4; ++cc) {
            int ei = i - 1 + (cc & 1), ek = k - 1 + ((cc >> 1) & 1);
            if (ei < 0 || ek < 0 || ei >= g.n0 - 1 || ek >= g.n1 - 1) continue;
            long long e = (long long)ek * (g.n0 - 1) + ei;
            double w = g.hx[ei], h = g.hy[ek];
            double r = g.rmid ? g.rmid[ei] : 1.;
            double kx1 = r * kh[e] * h / w, ky1 = r * kv[e] * w / h;
            b += (kx1 + ky1) / 3.;
            if (ek == k) c += (kx1 - 2. * ky1) / 6.;
        }
        if (g.dsh) b += g.dsh[node];
        const bool dk = dbit(dmask, node);
        if (dk) b = 1.;
        if (dk || k == g.n1 - 1 || dbit(dmask, node + g.n0)) c = 0.;
    }
};

// Whole-line Thomas factors (one thread per column, sequential in k).
template <typename T, class BC>
__device__ void line_setup_column(const BC& bc, int n, long long plane, int col, T* __restrict__ la,
                                  T* __restrict__ lw, T* __restrict__ lcp) {
    double cprev = 0., cpprev = 0.;
    for (int k = 0; k < n; ++k) {
        const long long node = (long long)k * plane + col;
        double b, c;
        bc(k, node, b, c);
        const double a = cprev;
        const double w = 1. / (b - a * cpprev);
        const double cp = c * w;
        la[node] = T(a);
        lw[node] = T(w);
        lcp[node] = T(cp);
        cprev = c;
        cpprev = cp;
    }
}

template <typename T>
__global__ void k_line_setup3d(Axes3<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                               const uint64_t* __restrict__ dmask, T* __restrict__ la, T* __restrict__ lw,
                               T* __restrict__ lcp) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.n0 * g.n1) return;
    const long long plane = (long long)g.n0 * g.n1;
    LineBC3 bc{g, kh, kv, dmask, col % g.n0, col / g.n0, plane};
    line_setup_column<T>(bc, g.n2, plane, col, la, lw, lcp);
}

template <typename T>
__global__ void k_line_setup2d(Axes2<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                               const uint64_t* __restrict__ dmask, T* __restrict__ la, T* __restrict__ lw,
                               T* __restrict__ lcp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n0) return;
    LineBC2 bc{g, kh, kv, dmask, i};
    line_setup_column<T>(bc, g.n1, g.n0, i, la, lw, lcp);
}

// ---------------------------------------------------------------------------
// Partitioned line solve (exact, same M as the whole-line Thomas).  Each line of
// n planes is cut into P = ceil(n / m) segments of m planes; the last plane of
// every segment but the final one is a separator.  With I = the interiors (the
// other planes) and s = the separators:
//   z_I = T_II^-1 (r_I - T_Is z_s),   S z_s = r_s - T_sI T_II^-1 r_I,
//   S = T_ss - T_sI T_II^-1 T_Is   (tridiagonal, SPD, size P - 1).
// T_II is block diagonal (one tridiagonal block per interior), so each segment
// is solved by its own thread with its planes held in registers; the small S system
// of each column is then solved from its Thomas factors (by a shuffle scan over the
// separators, or one thread's walk for lines of more than 64 separators; part_solve).
// Stored per node:
//   interior plane: la = coupling to the previous interior plane (0 on the
//     first), lw = LDL^T pivot inverse, lcp = c' (0 on the last);
//   separator: la = coupling to the plane below, lcp = coupling to the plane
//     above (lw holds the diagonal, read only here).
// schur[(0|1|2)(P-1) + j][col] = Thomas factors (a, w, c') of S.
// S entries from the interior factors: (T_j^-1)_{LL} = w_L (the last pivot
// inverse); with v = L^-1 e_F (v_0 = 1, v_k = -la_k w_{k-1} v_{k-1}):
// (T^-1)_{FF} = sum_k v_k^2 w_k and (T^-1)_{LF} = w_L v_L.
// ---------------------------------------------------------------------------
// The factors are set up in three parallel steps (a serial walk up each column costs
// ~1 ms per grid whatever its size, as the planes are the critical path):
// 1. per node: the line's diagonal b and upper coupling c (LineBC3 / LineBC2);
// 2. per (column, segment): the interior factors of the segment, and the two sums of the segment's inverse the separator step needs
//    (ff = e_1^T T^-1 e_1 and fl = its first-to-last entry, from the stored factors);
// 3. per column: the LU recurrence of the separator (Schur complement) system.
template <typename BC>
__device__ __forceinline__ void line_bc_node(const BC& bc, int k, long long node, double* __restrict__ bb,
                                             double* __restrict__ cc) {
    double b, c;
    bc(k, node, b, c);
    bb[node] = b;
    cc[node] = c;
}

__global__ void k_line_bc3d(Axes3<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                            const uint64_t* __restrict__ dmask, double* __restrict__ bb, double* __restrict__ cc) {
    const long long plane = (long long)g.n0 * g.n1, N = plane * g.n2;
    for (long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        const int col = (int)(node % plane);
        const LineBC3 bc{g, kh, kv, dmask, col % g.n0, col / g.n0, plane};
        line_bc_node(bc, (int)(node / plane), node, bb, cc);
    }
}

__global__ void k_line_bc2d(Axes2<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                            const uint64_t* __restrict__ dmask, double* __restrict__ bb, double* __restrict__ cc) {
    const long long N = (long long)g.n0 * g.n1;
    for (long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        const LineBC2 bc{g, kh, kv, dmask, (int)(node % g.n0)};
        line_bc_node(bc, (int)(node / g.n0), node, bb, cc);
    }
}

// thread = (column, segment); sff / sfl: [segment][column] (segment 0 unused)
template <typename T>
__global__ void k_part_factor(int ncol, int n, int m, long long plane, const double* __restrict__ bb,
                              const double* __restrict__ cc, T* __restrict__ la, T* __restrict__ lw,
                              T* __restrict__ lcp, double* __restrict__ sff, double* __restrict__ sfl) {
    const int P = (n + m - 1) / m;
    const long long total = (long long)P * ncol;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int col = (int)(t % ncol), s = (int)(t / ncol);
        const int k0 = s * m, k1 = min(k0 + m, n);
        double cprev = 0., cpprev = 0.;
        for (int k = k0; k < k1; ++k) {
            const long long node = (long long)k * plane + col;
            const double b = bb[node], c = cc[node];
            const int u = k - k0;
            if (u == m - 1 && s < P - 1) {  // separator
                la[node] = T(cprev);
                lw[node] = T(b);
                lcp[node] = T(c);
            } else {
                const bool last = s < P - 1 ? u == m - 2 : k == n - 1;
                const double a = u == 0 ? 0. : cprev;
                const double w = 1. / (b - a * cpprev);
                const double cp = last ? 0. : c * w;
                la[node] = T(a);
                lw[node] = T(w);
                lcp[node] = T(cp);
                cprev = c;
                cpprev = cp;
            }
        }
        if (s == 0) continue;
        // interior of this segment (what follows the separator of segment s - 1), from
        // the stored (precision T) factors as the separator step reads them
        const int ke = s < P - 1 ? k0 + m - 1 : n;
        double v = 1., ff = 0., wprev = 0.;
        for (int k = k0; k < ke; ++k) {
            const long long node = (long long)k * plane + col;
            const double w = lw[node];
            if (k > k0) v = -(double)la[node] * wprev * v;
            ff += v * v * w;
            wprev = w;
        }
        sff[(size_t)s * ncol + col] = ff;
        sfl[(size_t)s * ncol + col] = v * wprev;
    }
}

template <typename T>
__global__ void k_part_schur(int ncol, int n, int m, long long plane, const T* __restrict__ la,
                             const T* __restrict__ lw, const T* __restrict__ lcp, const double* __restrict__ sff,
                             const double* __restrict__ sfl, T* __restrict__ schur) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int P = (n + m - 1) / m;
    double off_prev = 0., sc_prev = 0.;
    for (int j = 0; j < P - 1; ++j) {
        const long long ns = (long long)((j + 1) * m - 1) * plane + col;
        const double elo = la[ns], bs = lw[ns], ehi = lcp[ns];
        const double wl = lw[ns - plane];
        const double ff = sff[(size_t)(j + 1) * ncol + col], fl = sfl[(size_t)(j + 1) * ncol + col];
        const double sjj = bs - elo * elo * wl - ehi * ehi * ff;
        const double soff = j + 1 < P - 1 ? -ehi * fl * (double)la[(long long)((j + 2) * m - 1) * plane + col] : 0.;
        const double w = 1. / (sjj - off_prev * sc_prev);
        const double cp = soff * w;
        schur[(size_t)j * ncol + col] = T(off_prev);
        schur[(size_t)(P - 1 + j) * ncol + col] = T(w);
        schur[(size_t)(2 * (P - 1) + j) * ncol + col] = T(cp);
        off_prev = soff;
        sc_prev = cp;
    }
}

// Column kernel: one thread per line segment (blockIdx.y selects the segment; with
// seg >= the line length a segment is the whole line).  START: z = M^-1 r only
// (beginning of a CG run); otherwise r -= alpha q first.  The forward sweep keeps d
// in z, the backward sweep overwrites it with the solution; (r, z) = r^T M^-1 r is
// summed in the forward sweep as sum d_k y_k with y = L^-1 r (M = L D L^T, d = D^-1 y),
// so r is not read again.  (x += alpha p is done by the next fused apply.)
template <typename T, bool START>
__global__ __launch_bounds__(64) void k_cg_line(int ncol, int nplanes_all, long long plane, int seg,
                                                T* __restrict__ r, const T* __restrict__ q, const T* __restrict__ la,
                                                const T* __restrict__ lw, const T* __restrict__ lcp,
                                                T* __restrict__ z, PcgState* __restrict__ st,
                                                double* __restrict__ partials) {
    if (!START && st->done) return;
    constexpr int U = 16;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = col < ncol;
    // this thread's line segment [k_begin, k_begin + nplanes) (blockIdx.y)
    const int k_begin = blockIdx.y * seg;
    const int nplanes = min(seg, nplanes_all - k_begin);
    r += (long long)k_begin * plane;
    if (!START) q += (long long)k_begin * plane;
    la += (long long)k_begin * plane;
    lw += (long long)k_begin * plane;
    lcp += (long long)k_begin * plane;
    z += (long long)k_begin * plane;
    const T alpha = START ? T(0) : T(st->alpha);
    double rr = 0., rz = 0.;
    T dprev = T(0);
    for (int k0 = 0; k0 < nplanes; k0 += U) {
        T rv[U], qv[U], av[U], wv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u;
            if (ok && k < nplanes) {
                const long long node = (long long)k * plane + col;
                rv[u] = r[node];
                av[u] = la[node];
                wv[u] = lw[node];
                if (!START) qv[u] = q[node];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u;
            if (ok && k < nplanes) {
                const long long node = (long long)k * plane + col;
                T ri = rv[u];
                if (!START) {
                    ri = ri - alpha * qv[u];
                    r[node] = ri;
                }
                rr += (double)ri * (double)ri;
                const T y = ri - av[u] * dprev;  // (L^-1 r)_k
                const T d = y * wv[u];
                rz += (double)d * (double)y;
                z[node] = d;
                dprev = d;
            }
        }
    }
    T znext = T(0);
    const int kl = nplanes - 1;
    for (int k0 = kl; k0 >= 0; k0 -= U) {
        T dv[U], cv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 - u;
            if (ok && k >= 0) {
                const long long node = (long long)k * plane + col;
                dv[u] = z[node];
                cv[u] = lcp[node];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 - u;
            if (ok && k >= 0) {
                const long long node = (long long)k * plane + col;
                const T zk = dv[u] - cv[u] * znext;
                z[node] = zk;
                znext = zk;
            }
        }
    }
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2>(v, partials, &st->counter[1], tot)) {
        if (START) finish_start(st, tot);
        else finish_update(st, tot);
    }
}

// Threads per block of the partitioned kernel: blocks of 1024 while the register
// arrays (5 m values) stay small, else 512.
template <typename T, int M>
struct PartLB {
    static constexpr int value = (M * (int)sizeof(T) <= 64) ? 1024 : 512;
};

// Placement of one thread in a partitioned line block: column blk * CB + cl (blk: the
// block's logical index),
// planes [seg m, seg m + m) (lanes run over columns, so every load is a contiguous
// row segment).  The last plane of every segment but the final one is a separator.
struct PartPos {
    int t, cl, seg, col, len, lint;
    bool act, sep;
    long long base;
};

__device__ inline PartPos part_pos(int ncol, int n, int P, int CB, long long plane, int m, int blk) {
    PartPos q;
    q.t = threadIdx.x;
    q.cl = q.t % CB;
    q.seg = q.t / CB;
    q.col = blk * CB + q.cl;
    q.act = q.seg < P && q.col < ncol;
    const int k0 = q.seg * m;
    q.len = q.act ? min(m, n - k0) : 0;
    q.sep = q.act && q.seg < P - 1;
    q.lint = q.sep ? q.len - 1 : q.len;
    q.base = (long long)k0 * plane + q.col;
    return q;
}

// Width of the shuffle scan that solves a column's separator system in part_solve (the
// power of two >= P - 1), or 0 where one thread walks each column instead: more than 64
// separators, or fewer threads in the block than CB * W.
__host__ __device__ inline int part_scan_width(int P, int CB, int S) {
    int W = 1;
    while (W < P - 1) W <<= 1;
    return (W <= 64 && CB * W <= S) ? W : 0;
}

// LDS values of part_solve for blocks of S threads
__host__ __device__ inline int part_lds_values(int P, int CB, int S) {
    return 6 * (S + (part_scan_width(P, CB, S) ? P : 0));
}

// z = T^-1 r for the thread's segment (values in registers): interior solve
// y = T_j^-1 r; the separator right-hand sides g = r_s - T_sI y go to LDS; S z_s = g is
// solved per column (a shuffle scan, or one thread); each interior is solved again with the separator
// values moved to its right-hand side.  Every thread of the block calls it (two
// barriers); lds holds part_lds_values() values: six arrays of P rows of CB values
// (segment-major), with one pad per row where the scan runs, so that the lanes of a
// column's scan, which step through the rows, fall in different banks.
template <typename T, int M>
__device__ inline void part_solve(const PartPos& q, int P, int CB, int ncol, const T* __restrict__ schur, T* lds,
                                  const T (&rv)[M], const T (&av)[M], const T (&wv)[M], const T (&cv)[M], T (&zv)[M]) {
    const int S = blockDim.x, W = part_scan_width(P, CB, S);
    const int CS = W ? CB + 1 : CB, SP = part_lds_values(P, CB, S) / 6;
    T* s_yf = lds;
    T* s_g = s_yf + SP;
    T* s_eh = s_g + SP;
    T* s_sa = s_eh + SP;
    T* s_sw = s_sa + SP;
    T* s_sc = s_sw + SP;
    const int t = q.t, cl = q.cl, seg = q.seg;
    const int pt = seg * CS + cl;  // the thread's slot (threads past CB * P have none)
    if (q.sep) {
        s_sa[pt] = schur[(size_t)seg * ncol + q.col];
        s_sw[pt] = schur[(size_t)(P - 1 + seg) * ncol + q.col];
        s_sc[pt] = schur[(size_t)(2 * (P - 1) + seg) * ncol + q.col];
    }
    // interior solve with the separators at zero
    T d[M];
    T dp = T(0);
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < q.lint) {
            const T y = rv[u] - av[u] * dp;
            d[u] = y * wv[u];
            dp = d[u];
        }
    }
    T zn = T(0), yl = T(0);
#pragma unroll
    for (int u = M - 1; u >= 0; --u) {
        if (u < q.lint) {
            const T zz = d[u] - cv[u] * zn;
            if (u == q.lint - 1) yl = zz;
            d[u] = zz;
            zn = zz;
        }
    }
    T rs = T(0), elo = T(0), ehi = T(0);
#pragma unroll
    for (int u = 0; u < M; ++u)
        if (q.sep && u == q.len - 1) {
            rs = rv[u];
            elo = av[u];
            ehi = cv[u];
        }
    if (seg < P) {
        s_yf[pt] = q.act ? d[0] : T(0);
        s_g[pt] = q.sep ? rs - elo * yl : T(0);
        s_eh[pt] = q.sep ? ehi : T(0);
    }
    __syncthreads();
    // The separator system S z_s = g of each column, S = L U from its Thomas factors
    // (L: pivots 1/w, sub-diagonal a; U: unit, super-diagonal c').  Up to 64 separators:
    // the lanes of W-wide groups (W the power of two >= P - 1; one group per column)
    // take one separator each and run both bidiagonal solves as affine first-order
    // recurrences x_j = a_j x_(j-1) + b_j, with a log2(W)-step scan over shuffles
    // ((a, b) after (a', b') composes to (a a', a b' + b)) instead of a walk of P - 1
    // dependent steps.  Longer lines walk each column with one thread.
    const int ns = P - 1;
    if (W) {
        const int j = t & (W - 1), c = t / W;
        const bool on = j < ns && c < CB && q.col - cl + c < ncol;
        const int o = j * CS + c;
        T a = T(0), b = T(0);
        if (on) {
            const T w = s_sw[o];
            a = -s_sa[o] * w;
            b = (s_g[o] - s_eh[o] * s_yf[o + CS]) * w;
        }
        for (int d = 1; d < W; d <<= 1) {  // L y = g
            const T ap = __shfl_up(a, d, W), bp = __shfl_up(b, d, W);
            if (j >= d) {
                b = fma(a, bp, b);
                a *= ap;
            }
        }
        a = on ? -s_sc[o] : T(0);
        if (!on) b = T(0);
        for (int d = 1; d < W; d <<= 1) {  // U z = y
            const T an = __shfl_down(a, d, W), bn = __shfl_down(b, d, W);
            if (j + d < W) {
                b = fma(a, bn, b);
                a *= an;
            }
        }
        if (on) s_g[o] = b;
    } else if (seg == 0 && q.col < ncol) {
        T dq = T(0);
        for (int j = 0; j < P - 1; ++j) {
            const int o = j * CS + cl;
            const T g = s_g[o] - s_eh[o] * s_yf[o + CS];
            const T dd = (g - s_sa[o] * dq) * s_sw[o];
            s_g[o] = dd;
            dq = dd;
        }
        T zq = T(0);
        for (int j = P - 2; j >= 0; --j) {
            const int o = j * CS + cl;
            const T zz = s_g[o] - s_sc[o] * zq;
            s_g[o] = zz;
            zq = zz;
        }
    }
    __syncthreads();
    const T lo = (q.act && seg > 0) ? s_g[pt - CS] : T(0);
    const T elp = (q.act && seg > 0) ? s_eh[pt - CS] : T(0);
    const T hi = q.sep ? s_g[pt] : T(0);
    // interior solve with the separator values on the right-hand side
    dp = T(0);
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < q.lint) {
            T rhs = rv[u];
            if (u == 0) rhs -= elp * lo;
            if (u == q.lint - 1) rhs -= elo * hi;
            const T y = rhs - av[u] * dp;
            d[u] = y * wv[u];
            dp = d[u];
        }
    }
    zn = T(0);
#pragma unroll
    for (int u = M - 1; u >= 0; --u) {
        if (u < q.lint) {
            const T zz = d[u] - cv[u] * zn;
            d[u] = zz;
            zn = zz;
        }
    }
#pragma unroll
    for (int u = 0; u < M; ++u) zv[u] = u < q.lint ? d[u] : hi;
}

// One CG update with the partitioned line solve: r -= alpha q (skipped for START),
// z = T^-1 r (part_solve), and (r, z), (r, r) reduced over the grid.
template <typename T, int M, bool START>
__global__ __launch_bounds__((PartLB<T, M>::value)) void k_cg_part(
    int ncol, int n, int P, int CB, long long plane, T* __restrict__ r, const T* __restrict__ q,
    const T* __restrict__ la, const T* __restrict__ lw, const T* __restrict__ lcp, const T* __restrict__ schur,
    T* __restrict__ z, PcgState* __restrict__ st, double* __restrict__ partials) {
    if (!START && st->done) return;
    extern __shared__ __align__(16) unsigned char part_smem[];
    const PartPos pp = part_pos(ncol, n, P, CB, plane, M, (int)blockIdx.x);
    T rv[M], av[M], wv[M], cv[M], qv[M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < pp.len) {
            const long long node = pp.base + (long long)u * plane;
            rv[u] = r[node];
            av[u] = la[node];
            wv[u] = lw[node];
            cv[u] = lcp[node];
            if (!START) qv[u] = q[node];
        }
    }
    const T alpha = START ? T(0) : T(st->alpha);
    double rr = 0., rz = 0.;
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < pp.len) {
            if (!START) {
                rv[u] = rv[u] - alpha * qv[u];
                r[pp.base + (long long)u * plane] = rv[u];
            }
            rr += (double)rv[u] * (double)rv[u];
        }
    }
    T zv[M];
    part_solve<T, M>(pp, P, CB, ncol, schur, reinterpret_cast<T*>(part_smem), rv, av, wv, cv, zv);
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < pp.len) {
            z[pp.base + (long long)u * plane] = zv[u];
            rz += (double)rv[u] * (double)zv[u];
        }
    }
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2>(v, partials, &st->counter[1], tot)) {
        if (START) finish_start(st, tot);
        else finish_update(st, tot);
    }
}

// Grid transfer fused into a multigrid sweep (template parameter XF of k_mg_smooth):
// XF 1, restriction: the sweep's right-hand side is R r_f of the finer level's residual
// src (computed per node and stored into r); XF 2, prolongation: the sweep starts from
// xin + P src, src the coarser level's correction.  dmask is the Dirichlet mask of the
// level the sweep works on; the tables are those between the two levels (mg_kernels.hpp).
template <typename T> struct MgXfer {
    int n0f = 0, n1f = 0, n0c = 0, n1c = 0;
    const int* cx = nullptr;
    const int* cy = nullptr;
    const int* mx = nullptr;
    const int* my = nullptr;
    const T* wx = nullptr;
    const T* wy = nullptr;
    const T* src = nullptr;
    const uint64_t* dmask = nullptr;
};

// the (up to) 3 fine nodes of coarse node I along one axis and their restriction
// weights (absent ones repeat the first index with weight 0)
template <typename T>
__device__ __forceinline__ void mg_restrict_axis(int I, int nc, const int* __restrict__ cidx,
                                                 const int* __restrict__ m, const T* __restrict__ w, int (&idx)[3],
                                                 T (&wt)[3]) {
    const int lo = I > 0 ? cidx[I - 1] + 1 : 0;
    const int hi = I < nc - 1 ? cidx[I + 1] - 1 : cidx[I];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = lo + a;
        const bool ok = i <= hi;
        idx[a] = ok ? i : lo;
        const int mm = m[idx[a]];
        const T ww = w[idx[a]];
        wt[a] = ok ? (mm == I ? T(1) - ww : (mm + 1 == I ? ww : T(0))) : T(0);
    }
}

// Damped line-Jacobi sweep of the multigrid smoother, same partitioned line solve:
// xout = xin + omega T^-1 r (xin null: start from zero; xin may be xout).
// UPD: the PCG residual update r -= alpha q is applied (and stored) first.
// FIN 1 / 2: this sweep yields the preconditioned PCG residual z = xout, and
// (rp, z), (rp, rp) of the PCG residual rp end the iteration (1) or start a run (2).
// XF: a fused grid transfer (MgXfer).  Exits at once after the stop (st->done).
template <typename T, int M, bool UPD, int FIN, int XF = 0>
__global__ __launch_bounds__((PartLB<T, M>::value)) void k_mg_smooth(
    int ncol, int n, int P, int CB, long long plane, T* __restrict__ r, const T* __restrict__ q, const T* xin,
    T* xout, const T* __restrict__ rp, T omega, const T* __restrict__ la, const T* __restrict__ lw,
    const T* __restrict__ lcp, const T* __restrict__ schur, PcgState* __restrict__ st,
    double* __restrict__ partials, MgXfer<T> xf, int xcd) {
    if (st->done) return;
    extern __shared__ __align__(16) unsigned char part_smem[];
    // XCD-aware order (xcd != 0, grid a multiple of 8): hardware block b runs on XCD b % 8;
    // each XCD takes a contiguous range of column groups, i.e. of the in-plane rows that
    // the level's residual tiles on the same XCD cover, so each kernel of a level finds the
    // vectors the previous one wrote in its own L2.  Blocks past the last group idle but
    // still join the grid reduction.
    const int blk = xcd ? (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const PartPos pp = part_pos(ncol, n, P, CB, plane, M, blk);
    T rv[M], av[M], wv[M], cv[M];
    const T alpha = UPD ? T(st->alpha) : T(0);
    // transfer tables of this thread's column (XF 1: its fine support; XF 2: its coarse cell)
    int ii[3] = {0, 0, 0}, jj[3] = {0, 0, 0};
    T wi[3] = {T(0), T(0), T(0)}, wj[3] = {T(0), T(0), T(0)};
    const int ccol = pp.act ? pp.col : 0;
    if constexpr (XF == 1) {
        mg_restrict_axis(ccol % xf.n0c, xf.n0c, xf.cx, xf.mx, xf.wx, ii, wi);
        mg_restrict_axis(ccol / xf.n0c, xf.n1c, xf.cy, xf.my, xf.wy, jj, wj);
    }
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < pp.len) {
            const long long node = pp.base + (long long)u * plane;
            if constexpr (XF == 1) {
                const long long k = (long long)pp.seg * M + u;
                const T* src = xf.src + k * (long long)xf.n0f * xf.n1f;
                T s = T(0);
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const T* row = src + (long long)jj[b] * xf.n0f;
                    s += wj[b] * (wi[0] * row[ii[0]] + wi[1] * row[ii[1]] + wi[2] * row[ii[2]]);
                }
                rv[u] = dbit(xf.dmask, node) ? T(0) : s;
                r[node] = rv[u];
            } else {
                rv[u] = r[node];
            }
            av[u] = la[node];
            wv[u] = lw[node];
            cv[u] = lcp[node];
            if (UPD) {
                rv[u] = rv[u] - alpha * q[node];
                r[node] = rv[u];
            }
        }
    }
    T zv[M];
    part_solve<T, M>(pp, P, CB, ncol, schur, reinterpret_cast<T*>(part_smem), rv, av, wv, cv, zv);
    double rz = 0., rr = 0.;
    MgProl<T> pr;
    if constexpr (XF == 2) {
        pr.n0c = xf.n0c;
        pr.n1c = xf.n1c;
        pr.mx = xf.mx;
        pr.my = xf.my;
        pr.wx = xf.wx;
        pr.wy = xf.wy;
        pr.xc = xf.src;
    }
#pragma unroll
    for (int u = 0; u < M; ++u) {
        if (u < pp.len) {
            const long long node = pp.base + (long long)u * plane;
            T x0 = xin ? xin[node] : T(0);
            if constexpr (XF == 2) {
                if (!dbit(xf.dmask, node))
                    x0 += mg_prol_at(pr, (long long)pp.seg * M + u, ccol % xf.n0f, ccol / xf.n0f);
            }
            const T xv = x0 + omega * zv[u];
            xout[node] = xv;
            if (FIN) {
                const double g = (double)(UPD ? rv[u] : rp[node]);
                rz += g * (double)xv;
                rr += g * g;
            }
        }
    }
    if (FIN) {
        double v[2] = {rz, rr}, tot[2];
        if (grid_reduce<2>(v, partials, &st->counter[1], tot)) {
            if (FIN == 2) finish_start(st, tot);
            else finish_update(st, tot);
        }
    }
}

// ---------------------------------------------------------------------------
// setup kernels (node-centric gathers over the elements touching a node)
// ---------------------------------------------------------------------------

// Dinv = 1 / diag(A) (1 at Dirichlet nodes); 3D diag = sum_e (kx'+ky'+kz')/9.
template <typename T>
__global__ void k_diag3d(Axes3<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                         const uint64_t* __restrict__ dmask, T* __restrict__ dinv) {
    const long long N = (long long)g.n0 * g.n1 * g.n2;
    for (long long node = blockIdx.x * (long long)blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        int i = node % g.n0;
        int j = (node / g.n0) % g.n1;
        int k = node / ((long long)g.n0 * g.n1);
        double d = 0.;
        for (int c = 0; c < 8; ++c) {
            int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1), ek = k - 1 + ((c >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= g.n0 - 1 || ej >= g.n1 - 1 || ek >= g.n2 - 1) continue;
            long long e = ((long long)ek * (g.n1 - 1) + ej) * (g.n0 - 1) + ei;
            double dx = g.hx[ei], dy = g.hy[ej], dz = g.hz[ek];
            double kxy = kh[e] * 1e-6, kz = kv[e] * 1e-6;
            d += (kxy * dy * dz / dx + kxy * dx * dz / dy + kz * dx * dy / dz) / 9.;
        }
        if (g.dsh) d += g.dsh[node];
        dinv[node] = dbit(dmask, node) ? T(1) : T(1. / d);
    }
}

template <typename T>
__global__ void k_diag2d(Axes2<double> g, const double* __restrict__ kh, const double* __restrict__ kv,
                         const uint64_t* __restrict__ dmask, T* __restrict__ dinv) {
    const long long N = (long long)g.n0 * g.n1;
    for (long long node = blockIdx.x * (long long)blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        int i = node % g.n0;
        int j = node / g.n0;
        double d = 0.;
        for (int c = 0; c < 4; ++c) {
            int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1);
            if (ei < 0 || ej < 0 || ei >= g.n0 - 1 || ej >= g.n1 - 1) continue;
            long long e = (long long)ej * (g.n0 - 1) + ei;
            double w = g.hx[ei], h = g.hy[ej];
            double r = g.rmid ? g.rmid[ei] : 1.;
            d += r * (kh[e] * h / w + kv[e] * w / h) / 3.;
        }
        if (g.dsh) d += g.dsh[node];
        dinv[node] = dbit(dmask, node) ? T(1) : T(1. / d);
    }
}

// Node load vector F: 3D sum_e 0.125e-18 dx dy dz q_e (therm3d.cpp:223)
__global__ void k_load3d(Axes3<double> g, const double* __restrict__ q, double* __restrict__ F) {
    const long long N = (long long)g.n0 * g.n1 * g.n2;
    for (long long node = blockIdx.x * (long long)blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        int i = node % g.n0;
        int j = (node / g.n0) % g.n1;
        int k = node / ((long long)g.n0 * g.n1);
        double f = 0.;
        for (int c = 0; c < 8; ++c) {
            int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1), ek = k - 1 + ((c >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= g.n0 - 1 || ej >= g.n1 - 1 || ek >= g.n2 - 1) continue;
            long long e = ((long long)ek * (g.n1 - 1) + ej) * (g.n0 - 1) + ei;
            f += 0.125e-18 * g.hx[ei] * g.hy[ej] * g.hz[ek] * q[e];
        }
        F[node] = f;
    }
}

// 2D: sum_e 0.25e-12 (r) w h q_e (therm2d.cpp:220, :346)
__global__ void k_load2d(Axes2<double> g, const double* __restrict__ q, double* __restrict__ F) {
    const long long N = (long long)g.n0 * g.n1;
    for (long long node = blockIdx.x * (long long)blockDim.x + threadIdx.x; node < N;
         node += (long long)gridDim.x * blockDim.x) {
        int i = node % g.n0;
        int j = node / g.n0;
        double f = 0.;
        for (int c = 0; c < 4; ++c) {
            int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1);
            if (ei < 0 || ej < 0 || ei >= g.n0 - 1 || ej >= g.n1 - 1) continue;
            long long e = (long long)ej * (g.n0 - 1) + ei;
            double r = g.rmid ? g.rmid[ei] : 1.;
            f += 0.25e-12 * r * g.hx[ei] * g.hy[ej] * q[e];
        }
        F[node] = f;
    }
}

// ---------------------------------------------------------------------------
// small vector kernels
// ---------------------------------------------------------------------------
template <typename TO, typename TI>
__global__ void k_convert(long long n, const TI* __restrict__ a, TO* __restrict__ b) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = TO(a[i]);
}

// x[node] = val for the Dirichlet list; optional second target
__global__ void k_scatter(long long m, const long long* __restrict__ nodes, const double* __restrict__ vals,
                          double* __restrict__ x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        x[nodes[i]] = vals[i];
}

// Linear (bi-/trilinear) interpolation of node values at arbitrary points: per axis the
// interval by bisection over the node coordinates (ax: axis 0, then 1, then 2), points
// outside the grid (or NaN) get `fill`.  pts: npts x dim, row-major.
__global__ void k_interp_nodes(const double* __restrict__ ax, int dim, int n0, int n1, int n2,
                               const double* __restrict__ x, long long npts, const double* __restrict__ pts,
                               double fill, double* __restrict__ out) {
    const int n[3] = {n0, n1, n2};
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npts;
         p += (long long)gridDim.x * blockDim.x) {
        int lo[3] = {0, 0, 0};
        double t[3] = {0., 0., 0.};
        bool inside = true;
        const double* A = ax;
        for (int a = 0; a < dim; ++a) {
            const double c = pts[p * dim + a];
            const int na = n[a];
            if (!(c >= A[0] && c <= A[na - 1])) inside = false;
            int l = 0, h = na - 1;
            while (h - l > 1) {
                const int mid = (l + h) >> 1;
                if (A[mid] <= c) l = mid;
                else h = mid;
            }
            lo[a] = l;
            t[a] = (c - A[l]) / (A[l + 1] - A[l]);
            A += na;
        }
        if (!inside) {
            out[p] = fill;
            continue;
        }
        const long long s1 = n0, s2 = (long long)n0 * n1;
        const long long base = lo[0] + s1 * lo[1] + (dim == 3 ? s2 * lo[2] : 0);
        double v = 0.;
        for (int c = 0; c < (1 << dim); ++c) {
            double w = 1.;
            long long off = 0;
            for (int a = 0; a < dim; ++a) {
                const int b = (c >> a) & 1;
                w *= b ? t[a] : 1. - t[a];
                off += b ? (a == 0 ? 1 : (a == 1 ? s1 : s2)) : 0;
            }
            v += w * x[base + off];
        }
        out[p] = v;
    }
}

// x[node] += val (nodes distinct)
__global__ void k_scatter_add(long long m, const long long* __restrict__ nodes, const double* __restrict__ vals,
                              double* __restrict__ x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        x[nodes[i]] += vals[i];
}

// x += s * (double) d
template <typename T>
__global__ void k_add_correction(long long n, double* __restrict__ x, const T* __restrict__ d, double s) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] += s * (double)d[i];
}

// b32 = (float)(r * s), zero initial guess handled by caller
__global__ void k_scale_to_f32(long long n, const double* __restrict__ r, double s, float* __restrict__ r32) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        r32[i] = (float)(r[i] * s);
}

template <typename T>
__global__ void k_fill(long long n, T* __restrict__ a, T v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = v;
}

// ---------------------------------------------------------------------------
// Picard-loop support: element means, k(T) laws, change statistics, junctions,
// currents and Joule heat.  Element loops are one thread per element.
// ---------------------------------------------------------------------------
struct ElemGrid {
    int dim, n0, n1, n2;  // nodes per axis (n2 = 1 in 2D)
};

__device__ __forceinline__ void elem_coords(const ElemGrid& g, long long e, int& i, int& j, int& k) {
    const int m0 = g.n0 - 1, m1 = g.n1 - 1;
    i = e % m0;
    if (g.dim == 2) {
        j = e / m0;
        k = 0;
    } else {
        j = (e / m0) % m1;
        k = e / ((long long)m0 * m1);
    }
}

// Heat flux -k grad T per element in W/m^2 (saveHeatFluxes, therm2d.cpp:504-540 and
// therm3d.cpp:342-384): the element's corner-difference gradient (x 0.5e6 in 2D, 0.25e6 in
// 3D: 1e6 um -> m over the corners averaged), in-plane k along x (and y), vertical k along
// the last axis.  k from the law k0 (300 / T_avg)^alpha at the element's mean temperature
// when ah is given, else from kh / kv.
__global__ void k_heat_flux(ElemGrid g, long long E, const double* __restrict__ hx, const double* __restrict__ hy,
                            const double* __restrict__ hz, const double* __restrict__ T, const double* __restrict__ kh,
                            const double* __restrict__ kv, const double* __restrict__ ah, const double* __restrict__ av,
                            double* __restrict__ flux) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_coords(g, e, i, j, k);
        if (g.dim == 2) {
            const long long ll = i + (long long)g.n0 * j, ul = ll + 1, lu = ll + g.n0, uu = lu + 1;
            const double tll = T[ll], tul = T[ul], tlu = T[lu], tuu = T[uu];
            double a = kh[e], b = kv[e];
            if (ah) {
                const double r = 300. / (0.25 * (tll + tul + tlu + tuu));
                a *= pow(r, ah[e]);
                b *= pow(r, av[e]);
            }
            flux[2 * e] = -0.5e6 * a * (-tll + tul - tlu + tuu) / hx[i];
            flux[2 * e + 1] = -0.5e6 * b * (-tll - tul + tlu + tuu) / hy[j];
        } else {
            double t[8], s = 0.;
            for (int c = 0; c < 8; ++c) {
                t[c] = T[(i + (c & 1)) + (long long)g.n0 * ((j + ((c >> 1) & 1)) + (long long)g.n1 * (k + (c >> 2)))];
                s += t[c];
            }
            double a = kh[e], b = kv[e];
            if (ah) {
                const double r = 300. / (0.125 * s);
                a *= pow(r, ah[e]);
                b *= pow(r, av[e]);
            }
            double dx = 0., dy = 0., dz = 0.;
            for (int c = 0; c < 8; ++c) {
                dx += (c & 1) ? t[c] : -t[c];
                dy += (c & 2) ? t[c] : -t[c];
                dz += (c & 4) ? t[c] : -t[c];
            }
            flux[3 * e] = -0.25e6 * a * dx / hx[i];
            flux[3 * e + 1] = -0.25e6 * a * dy / hy[j];
            flux[3 * e + 2] = -0.25e6 * b * dz / hz[k];
        }
    }
}

// Element (midpoint) field at points: linear between midpoints, the edge value between
// the outermost midpoint and the mesh boundary, fill outside the mesh.  ax holds the node
// axes, mid the midpoint axes (n - 1 each), both concatenated; nc components per element.
__global__ void k_interp_elems(const double* __restrict__ ax, const double* __restrict__ mid, int dim, int n0, int n1,
                               int n2, const double* __restrict__ f, int nc, long long npts,
                               const double* __restrict__ pts, double fill, double* __restrict__ out) {
    const int n[3] = {n0, n1, n2};
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npts;
         p += (long long)gridDim.x * blockDim.x) {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        double t[3] = {0., 0., 0.};
        bool inside = true;
        const double* A = ax;
        const double* M = mid;
        for (int a = 0; a < dim; ++a) {
            const double c = pts[p * dim + a];
            const int na = n[a], m = na - 1;
            if (!(c >= A[0] && c <= A[na - 1])) inside = false;
            const double cc = fmin(fmax(c, M[0]), M[m - 1]);
            int l = 0, h = m - 1;
            while (h - l > 1) {
                const int md = (l + h) >> 1;
                if (M[md] <= cc) l = md;
                else h = md;
            }
            lo[a] = l;
            hi[a] = m > 1 ? l + 1 : l;
            t[a] = m > 1 ? (cc - M[l]) / (M[l + 1] - M[l]) : 0.;
            A += na;
            M += m;
        }
        if (!inside) {
            for (int q = 0; q < nc; ++q) out[p * nc + q] = fill;
            continue;
        }
        const long long m0 = n0 - 1, m1 = n1 - 1;
        for (int q = 0; q < nc; ++q) {
            double v = 0.;
            for (int c = 0; c < (1 << dim); ++c) {
                double w = 1.;
                int id[3] = {0, 0, 0};
                for (int a = 0; a < dim; ++a) {
                    const int b = (c >> a) & 1;
                    w *= b ? t[a] : 1. - t[a];
                    id[a] = b ? hi[a] : lo[a];
                }
                const long long e = id[0] + m0 * (id[1] + (dim == 3 ? m1 * id[2] : 0));
                v += w * f[e * nc + q];
            }
            out[p * nc + q] = v;
        }
    }
}


__device__ __forceinline__ long long node_of(const ElemGrid& g, int i, int j, int k) {
    return i + (long long)g.n0 * (j + (long long)g.n1 * k);
}

// mean of the element's corner values (the element temperature of therm3d.cpp:209,
// therm2d.cpp:210)
__global__ void k_elem_mean(ElemGrid g, long long E, const double* __restrict__ x, double* __restrict__ out) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_coords(g, e, i, j, k);
        double s = 0.;
        const int nc = g.dim == 3 ? 8 : 4;
        for (int c = 0; c < nc; ++c) s += x[node_of(g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))];
        out[e] = s / nc;
    }
}

// k = k0 (300 / T_elem)^alpha for both tensor components; writes FP64 and FP32 copies
__global__ void k_powerlaw(ElemGrid g, long long E, const double* __restrict__ x, const double* __restrict__ k0h,
                           const double* __restrict__ k0v, const double* __restrict__ ah, const double* __restrict__ av,
                           double* __restrict__ kh, double* __restrict__ kv, float* __restrict__ kh32,
                           float* __restrict__ kv32) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        if (isnan(k0h[e])) continue;  // the element keeps its conductivities
        int i, j, k;
        elem_coords(g, e, i, j, k);
        double s = 0.;
        const int nc = g.dim == 3 ? 8 : 4;
        for (int c = 0; c < nc; ++c) s += x[node_of(g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))];
        const double t = 300. / (s / nc);
        const double h = ah[e] == 0. ? k0h[e] : k0h[e] * pow(t, ah[e]);
        const double v = av[e] == 0. ? k0v[e] : k0v[e] * pow(t, av[e]);
        kh[e] = h;
        kv[e] = v;
        kh32[e] = (float)h;
        kv32[e] = (float)v;
    }
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// max over the grid of NV values; last block returns totals
template <int NV>
__device__ bool grid_max(double (&v)[NV], double* __restrict__ partials, unsigned int* counter, double (&tot)[NV]) {
    __shared__ double red[NV][16];
    __shared__ bool last;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nw = (nthr + 63) >> 6;
    for (int a = 0; a < NV; ++a) {
        double s = wave_max(v[a]);
        if (lane == 0) red[a][wid] = s;
    }
    __syncthreads();
    const unsigned int nblk = gridDim.x, bid = blockIdx.x;
    const double ninf = -__builtin_huge_val();
    if (tid == 0) {
        for (int a = 0; a < NV; ++a) {
            double s = ninf;
            for (int w = 0; w < nw; ++w) s = fmax(s, red[a][w]);
            publish_partial(partials + (size_t)a * nblk + bid, s);
        }
        last = (take_ticket(counter) == nblk - 1);
    }
    __syncthreads();
    if (!last) return false;
    for (int a = 0; a < NV; ++a) {
        double s = ninf;
        for (unsigned int b = tid; b < nblk; b += nthr) s = fmax(s, read_partial(partials + (size_t)a * nblk + b));
        s = wave_max(s);
        __syncthreads();
        if (lane == 0) red[a][wid] = s;
        __syncthreads();
        if (tid == 0) {
            double t = ninf;
            for (int w = 0; w < nw; ++w) t = fmax(t, red[a][w]);
            tot[a] = t;
        }
    }
    if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return tid == 0;
}

// max |x - xprev| and max x (therm2d.cpp:480-487); then xprev = x.  Results in out[0..1].
__global__ __launch_bounds__(256) void k_delta_stats(long long n, const double* __restrict__ x, double* __restrict__ xprev,
                                                     double* __restrict__ out, double* __restrict__ partials,
                                                     unsigned int* counter) {
    double d = 0., m = -__builtin_huge_val();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double xi = x[i];
        d = fmax(d, fabs(xi - xprev[i]));
        m = fmax(m, xi);
        xprev[i] = xi;
    }
    double v[2] = {d, m}, tot[2];
    if (grid_max<2>(v, partials, counter, tot)) {
        out[0] = tot[0];
        out[1] = tot[1];
    }
}

// Junction conductivity update (electr2d.cpp:243-264, electr3d.cpp:246-273) with the
// law of BetaSolver::activeCond (beta.hpp:43-46).  For junction element q:
// U = mean_c (V[top_c] - V[bot_c]), jy = 0.1 sigma_yy U / d,
// sigma = 10 |jy| d beta / ln(1e7 |jy| / js + 1); STABLE averages with the old value;
// values below 1e-16 or NaN are clamped to 1e-16.  In-plane conductivity is 0.
__global__ void k_junction_update(long long nje, int ncorner, const long long* __restrict__ elems,
                                  const long long* __restrict__ bot, const long long* __restrict__ top,
                                  const int* __restrict__ jid, const double* __restrict__ height,
                                  const double* __restrict__ beta, const double* __restrict__ js, int stable,
                                  const double* __restrict__ V, double* __restrict__ kh, double* __restrict__ kv,
                                  float* __restrict__ kh32, float* __restrict__ kv32) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nje; q += (long long)gridDim.x * blockDim.x) {
        double U = 0.;
        for (int c = 0; c < ncorner; ++c) U += V[top[q * ncorner + c]] - V[bot[q * ncorner + c]];
        U /= ncorner;
        const long long e = elems[q];
        const int n = jid[q];
        const double d = height[n];
        double jy = fabs(0.1 * kv[e] * U / d);
        double cond = 10. * jy * d * beta[q] / log(1e7 * jy / js[q] + 1.);
        if (stable) cond = 0.5 * (kv[e] + cond);
        if (isnan(cond) || fabs(cond) < 1e-16) cond = 1e-16;
        // the law's in-plane component is 0 (averaged with the previous one when stable)
        const double ch = stable ? 0.5 * kh[e] : 0.;
        kv[e] = cond;
        kh[e] = ch;
        kv32[e] = (float)cond;
        kh32[e] = (float)ch;
    }
}

// U and jy of every junction element for a law evaluated by the caller
__global__ void k_junction_state(long long nje, int ncorner, const long long* __restrict__ elems,
                                 const long long* __restrict__ bot, const long long* __restrict__ top,
                                 const int* __restrict__ jid, const double* __restrict__ height,
                                 const double* __restrict__ V, const double* __restrict__ kv, double* __restrict__ Uo,
                                 double* __restrict__ jo) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nje; q += (long long)gridDim.x * blockDim.x) {
        double U = 0.;
        for (int c = 0; c < ncorner; ++c) U += V[top[q * ncorner + c]] - V[bot[q * ncorner + c]];
        U /= ncorner;
        Uo[q] = U;
        jo[q] = 0.1 * kv[elems[q]] * U / height[jid[q]];
    }
}

// junction conductivities from the caller's law
__global__ void k_junction_cond_set(long long nje, const long long* __restrict__ elems, const double* __restrict__ ch_in,
                                    const double* __restrict__ cv_in, int stable, double* __restrict__ kh,
                                    double* __restrict__ kv, float* __restrict__ kh32, float* __restrict__ kv32) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nje; q += (long long)gridDim.x * blockDim.x) {
        const long long e = elems[q];
        double ch = ch_in[q], cv = cv_in[q];
        if (stable) {
            ch = 0.5 * (kh[e] + ch);
            cv = 0.5 * (kv[e] + cv);
        }
        if (isnan(cv) || fabs(cv) < 1e-16) cv = 1e-16;
        kh[e] = ch;
        kv[e] = cv;
        kh32[e] = (float)ch;
        kv32[e] = (float)cv;
    }
}

// Element current densities j = -sigma grad V in kA/cm^2 (electr2d.cpp:405-409,
// electr3d.cpp:399-411), Joule heat sigma |grad V|^2 in W/m^3 (electr2d.cpp:457-461,
// electr3d.cpp:459-475; zero where noheat[e]), and the Picard error inputs:
// max |j - j_prev|^2 over all elements and max |j|^2 over the selected elements.
__global__ __launch_bounds__(256) void k_currents(ElemGrid g, long long E, const double* __restrict__ hx,
                                                  const double* __restrict__ hy, const double* __restrict__ hz,
                                                  const double* __restrict__ V, const double* __restrict__ kh,
                                                  const double* __restrict__ kv, const unsigned char* __restrict__ sel,
                                                  const unsigned char* __restrict__ noheat, double* __restrict__ cur,
                                                  double* __restrict__ heat, double* __restrict__ out,
                                                  double* __restrict__ partials, unsigned int* counter) {
    double dmax = 0., jmax = 0.;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_coords(g, e, i, j, k);
        double c[3] = {0., 0., 0.}, h = 0.;
        if (g.dim == 2) {
            double ll = V[node_of(g, i, j, 0)], lu = V[node_of(g, i + 1, j, 0)], ul = V[node_of(g, i, j + 1, 0)],
                   uu = V[node_of(g, i + 1, j + 1, 0)];
            double gx = (-ll + lu - ul + uu) / hx[i], gy = (-ll - lu + ul + uu) / hy[j];
            c[0] = kh[e] * (-0.05 * gx);
            c[1] = kv[e] * (-0.05 * gy);
            double ex = 0.5e6 * gx, ey = 0.5e6 * gy;
            h = kh[e] * ex * ex + kv[e] * ey * ey;
        } else {
            double sx = 0., sy = 0., sz = 0.;
            for (int a = 0; a < 8; ++a) {
                double v = V[node_of(g, i + (a & 1), j + ((a >> 1) & 1), k + ((a >> 2) & 1))];
                sx += (a & 1) ? v : -v;
                sy += (a & 2) ? v : -v;
                sz += (a & 4) ? v : -v;
            }
            c[0] = -0.025 * kh[e] * sx / hx[i];
            c[1] = -0.025 * kh[e] * sy / hy[j];
            c[2] = -0.025 * kv[e] * sz / hz[k];
            double ex = -0.25e6 * sx / hx[i], ey = -0.25e6 * sy / hy[j], ez = -0.25e6 * sz / hz[k];
            h = kh[e] * ex * ex + kh[e] * ey * ey + kv[e] * ez * ez;
        }
        double dd = 0., jj = 0.;
        for (int a = 0; a < g.dim; ++a) {
            double o = cur[e * g.dim + a];
            dd += (c[a] - o) * (c[a] - o);
            jj += c[a] * c[a];
            cur[e * g.dim + a] = c[a];
        }
        heat[e] = (noheat && noheat[e]) ? 0. : h;
        dmax = fmax(dmax, dd);
        if (!sel || sel[e]) jmax = fmax(jmax, jj);
    }
    double v[2] = {dmax, jmax}, tot[2];
    if (grid_max<2>(v, partials, counter, tot)) {
        out[0] = tot[0];
        out[1] = tot[1];
    }
}

}  // namespace pf
