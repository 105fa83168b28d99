// Grid-transfer and coarse-level set-up kernels of the semicoarsening multigrid
// preconditioner.  The coarse levels coarsen the in-plane axes only (every other node,
// plus the last one; in 2D the one in-plane axis 0, with a trivial one-node axis 1 in
// the tables below); the layer axis keeps all its planes, so every level has the same
// column structure and the exact line solve along the layer axis is the smoother on
// every level.
//
// Transfer tables per in-plane axis, from a finer level to the next coarser one:
//   cidx[I]          fine index of coarse node I (increasing, first 0, last n_f - 1)
//   tm[i], tw[i]     fine node i lies in coarse interval [tm[i], tm[i] + 1] at relative
//                    position tw[i] in [0, 1) (linear interpolation weight of the right
//                    node; 0 for nodes that are coarse nodes)
// Prolongation P = I_z (x) P_y (x) P_x is the Q1 interpolation of the coarse field at the
// fine nodes; restriction is its transpose; both skip Dirichlet nodes (the coarse
// correction is zero there).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fem_kernels.hpp"

namespace pf {

struct MgTransfer {
    int n0f, n1f, n0c, n1c, n2;
    const int* cx;  // [n0c]
    const int* cy;  // [n1c]
    const int* mx;  // [n0f]
    const int* my;  // [n1f]
    const int* rx;  // [3 n0c] restriction rows: fine indices per coarse node
    const int* ry;  // [3 n1c]
};

// weight of fine node i (left coarse index m, position w) in coarse node I
template <typename T>
__device__ __forceinline__ T tweight(int m, T w, int I) {
    return m == I ? T(1) - w : (m + 1 == I ? w : T(0));
}

// the (up to) 3 fine nodes of coarse node I along one axis and their weights: the
// fine nodes strictly between the neighbouring coarse nodes; absent ones repeat the
// first index with weight 0, so the 3 x 3 gather below has fixed addresses and all its
// loads issue together
template <typename T>
__device__ __forceinline__ void restrict_axis(int I, int nc, const int* __restrict__ cidx, const int* __restrict__ m,
                                              const T* __restrict__ w, int (&idx)[3], T (&wt)[3]) {
    const int lo = I > 0 ? cidx[I - 1] + 1 : 0;
    const int hi = I < nc - 1 ? cidx[I + 1] - 1 : cidx[I];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = lo + a;
        const bool ok = i <= hi;
        idx[a] = ok ? i : lo;
        wt[a] = ok ? tweight(m[idx[a]], w[idx[a]], I) : T(0);
    }
}

// (i, j, k) of node idx of an n0 x n1 x planes grid: 32-bit divisions below 2^31 nodes
// (the 64-bit ones are long software sequences on the critical path of these short
// kernels)
__device__ __forceinline__ void split_node(long long idx, long long N, int n0, int n1, int& i, int& j, long long& k) {
    if (N <= 0x7fffffffLL) {
        const unsigned u = (unsigned)idx, row = u / (unsigned)n0, kk = row / (unsigned)n1;
        i = (int)(u - row * (unsigned)n0);
        j = (int)(row - kk * (unsigned)n1);
        k = kk;
    } else {
        i = (int)(idx % n0);
        j = (int)((idx / n0) % n1);
        k = idx / ((long long)n0 * n1);
    }
}

// b_c = R r_f (transpose of the interpolation), zero at coarse Dirichlet nodes; a
// coarse node's support has at most 3 fine nodes per in-plane axis (coarsening by 2)
template <typename T>
__global__ __launch_bounds__(256) void k_mg_restrict(MgTransfer tr, const T* __restrict__ rwx,
                                                     const T* __restrict__ rwy, const T* __restrict__ rf,
                                                     T* __restrict__ bc, const uint64_t* __restrict__ dmask_c,
                                                     const PcgState* __restrict__ st) {
    if (st->done) return;
    const long long planec = (long long)tr.n0c * tr.n1c, planef = (long long)tr.n0f * tr.n1f;
    const long long Nc = planec * tr.n2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nc;
         idx += (long long)gridDim.x * blockDim.x) {
        int I, J;
        long long k;
        split_node(idx, Nc, tr.n0c, tr.n1c, I, J, k);
        // the coarse node's restriction rows (the same as restrict_axis gives, tabulated)
        int ii[3], jj[3];
        T wi[3], wj[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ii[a] = tr.rx[3 * I + a];
            wi[a] = rwx[3 * I + a];
            jj[a] = tr.ry[3 * J + a];
            wj[a] = rwy[3 * J + a];
        }
        const T* src = rf + k * planef;
        T s = T(0);
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T* row = src + (long long)jj[b] * tr.n0f;
            s += wj[b] * (wi[0] * row[ii[0]] + wi[1] * row[ii[1]] + wi[2] * row[ii[2]]);
        }
        bc[idx] = dbit(dmask_c, idx) ? T(0) : s;
    }
}

// x_f += P x_c at the free fine nodes
template <typename T>
__global__ __launch_bounds__(256) void k_mg_prolong_add(MgTransfer tr, const T* __restrict__ wx,
                                                        const T* __restrict__ wy, const T* __restrict__ xc,
                                                        T* __restrict__ xf, const uint64_t* __restrict__ dmask_f,
                                                        const PcgState* __restrict__ st) {
    if (st->done) return;
    const long long planec = (long long)tr.n0c * tr.n1c, planef = (long long)tr.n0f * tr.n1f;
    const long long Nf = planef * tr.n2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Nf;
         idx += (long long)gridDim.x * blockDim.x) {
        // the Dirichlet test selects the value instead of skipping the node, so the
        // mask, table and vector loads do not wait on a branch
        const bool fixed = dbit(dmask_f, idx);
        int i, j;
        long long k;
        split_node(idx, Nf, tr.n0f, tr.n1f, i, j, k);
        const int I = tr.mx[i], J = tr.my[j];
        const int I1 = min(I + 1, tr.n0c - 1), J1 = min(J + 1, tr.n1c - 1);
        const T a = wx[i], b = wy[j];
        const T xo = xf[idx];
        const T* c = xc + k * planec;
        const T lo = (T(1) - a) * c[(long long)J * tr.n0c + I] + a * c[(long long)J * tr.n0c + I1];
        const T hi = (T(1) - a) * c[(long long)J1 * tr.n0c + I] + a * c[(long long)J1 * tr.n0c + I1];
        xf[idx] = fixed ? xo : xo + ((T(1) - b) * lo + b * hi);
    }
}

// The same for an even number of fine nodes along axis 0, two nodes (i, i + 1), i even, per
// thread: their interval indices differ by 0 or 1, so the pair needs three coarse values
// per coarse row instead of four per node, and the pair's table, mask and vector accesses
// are single (8- or 16-byte) loads.  The interpolation formula per node is that of
// k_mg_prolong_add (results agree to rounding: the compiler may contract it into FMAs
// differently).
template <typename T>
struct alignas(2 * sizeof(T)) PairOf {
    T a, b;
};

template <typename T>
__global__ __launch_bounds__(256) void k_mg_prolong_add2(MgTransfer tr, const T* __restrict__ wx,
                                                         const T* __restrict__ wy, const T* __restrict__ xc,
                                                         T* __restrict__ xf, const uint64_t* __restrict__ dmask_f,
                                                         const PcgState* __restrict__ st) {
    if (st->done) return;
    const long long planec = (long long)tr.n0c * tr.n1c;
    const int h0 = tr.n0f / 2;  // pairs per row (n0f even)
    const long long Np = (long long)h0 * tr.n1f * tr.n2;
    for (long long pid = (long long)blockIdx.x * blockDim.x + threadIdx.x; pid < Np;
         pid += (long long)gridDim.x * blockDim.x) {
        int ip, j;
        long long k;
        split_node(pid, Np, h0, tr.n1f, ip, j, k);
        const int i = 2 * ip;
        const long long idx = 2 * pid;  // the pair's first fine node (rows hold whole pairs)
        const uint64_t w = dmask_f[idx >> 6];  // idx even: both nodes in one mask word
        const bool fixed0 = (w >> (idx & 63)) & 1ull, fixed1 = (w >> ((idx + 1) & 63)) & 1ull;
        const int2 Ip = *reinterpret_cast<const int2*>(tr.mx + i);
        const PairOf<T> ap = *reinterpret_cast<const PairOf<T>*>(wx + i);
        const PairOf<T> xo = *reinterpret_cast<const PairOf<T>*>(xf + idx);
        const int J = tr.my[j], J1 = min(J + 1, tr.n1c - 1);
        const T b = wy[j];
        const int I0 = Ip.x, I0n = min(I0 + 1, tr.n0c - 1), I1n = min(Ip.y + 1, tr.n0c - 1);
        const T* r0 = xc + k * planec + (long long)J * tr.n0c;
        const T* r1 = xc + k * planec + (long long)J1 * tr.n0c;
        const T u0 = r0[I0], u1 = r0[I0n], u2 = r0[I1n];
        const T v0 = r1[I0], v1 = r1[I0n], v2 = r1[I1n];
        const bool same = Ip.y == I0;  // else Ip.y == I0 + 1
        const T lo0 = (T(1) - ap.a) * u0 + ap.a * u1, hi0 = (T(1) - ap.a) * v0 + ap.a * v1;
        const T lo1 = (T(1) - ap.b) * (same ? u0 : u1) + ap.b * (same ? u1 : u2);
        const T hi1 = (T(1) - ap.b) * (same ? v0 : v1) + ap.b * (same ? v1 : v2);
        PairOf<T> out;
        out.a = fixed0 ? xo.a : xo.a + ((T(1) - b) * lo0 + b * hi0);
        out.b = fixed1 ? xo.b : xo.b + ((T(1) - b) * lo1 + b * hi1);
        *reinterpret_cast<PairOf<T>*>(xf + idx) = out;
    }
}

// Coarse element conductivities: coarse element (I, J, k) covers the fine elements
// [cx[I], cx[I+1]) x [cy[J], cy[J+1]) of element layer k; area-weighted means.
__global__ __launch_bounds__(256) void k_mg_coarsen_k(MgTransfer tr, const double* __restrict__ hxf,
                                                      const double* __restrict__ hyf, const double* __restrict__ khf,
                                                      const double* __restrict__ kvf, double* __restrict__ khc,
                                                      double* __restrict__ kvc) {
    const int e0c = tr.n0c - 1, e1c = tr.n1c - 1, e0f = tr.n0f - 1, e1f = tr.n1f - 1;
    const long long Ec = (long long)e0c * e1c * (tr.n2 - 1);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < Ec; e += (long long)gridDim.x * blockDim.x) {
        const int I = (int)(e % e0c);
        const int J = (int)((e / e0c) % e1c);
        const long long k = e / ((long long)e0c * e1c);
        const double* kh = khf + k * e0f * e1f;
        const double* kv = kvf + k * e0f * e1f;
        double sa = 0., sh = 0., sv = 0.;
        for (int ey = tr.cy[J]; ey < tr.cy[J + 1]; ++ey)
            for (int ex = tr.cx[I]; ex < tr.cx[I + 1]; ++ex) {
                const double a = hxf[ex] * hyf[ey];
                const long long f = (long long)ey * e0f + ex;
                sa += a;
                sh += a * kh[f];
                sv += a * kv[f];
            }
        khc[e] = sh / sa;
        kvc[e] = sv / sa;
    }
}

// 2D: coarse element (I, k) covers the fine elements [cx[I], cx[I+1]) of element row k;
// means weighted by the element width (Cartesian) or width x mid radius (cylindrical,
// the weight the element matrix carries).
__global__ __launch_bounds__(256) void k_mg_coarsen_k2d(MgTransfer tr, const double* __restrict__ hxf,
                                                        const double* __restrict__ rmidf,
                                                        const double* __restrict__ khf,
                                                        const double* __restrict__ kvf, double* __restrict__ khc,
                                                        double* __restrict__ kvc) {
    const int e0c = tr.n0c - 1, e0f = tr.n0f - 1;
    const long long Ec = (long long)e0c * (tr.n2 - 1);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < Ec; e += (long long)gridDim.x * blockDim.x) {
        const int I = (int)(e % e0c);
        const long long k = e / e0c;
        double sa = 0., sh = 0., sv = 0.;
        for (int ex = tr.cx[I]; ex < tr.cx[I + 1]; ++ex) {
            const double a = hxf[ex] * (rmidf ? rmidf[ex] : 1.);
            const long long f = k * e0f + ex;
            sa += a;
            sh += a * khf[f];
            sv += a * kvf[f];
        }
        khc[e] = sh / sa;
        kvc[e] = sv / sa;
    }
}

// Coarse Dirichlet mask: a coarse node is fixed when the fine node at its position is.
__global__ __launch_bounds__(256) void k_mg_coarsen_mask(MgTransfer tr, const uint64_t* __restrict__ dmf,
                                                         uint64_t* __restrict__ dmc) {
    const long long planec = (long long)tr.n0c * tr.n1c, planef = (long long)tr.n0f * tr.n1f;
    const long long Nc = planec * tr.n2;
    const long long words = (Nc + 63) / 64;
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < words;
         w += (long long)gridDim.x * blockDim.x) {
        uint64_t bits = 0;
        for (int b = 0; b < 64; ++b) {
            const long long idx = w * 64 + b;
            if (idx >= Nc) break;
            const int I = (int)(idx % tr.n0c);
            const int J = (int)((idx / tr.n0c) % tr.n1c);
            const long long k = idx / planec;
            const long long f = k * planef + (long long)tr.cy[J] * tr.n0f + tr.cx[I];
            if (dbit(dmf, f)) bits |= (uint64_t)1 << b;
        }
        dmc[w] = bits;
    }
}

}  // namespace pf
