// Persistent batched solver for many uneven 2D (Cartesian or cylindrical) Shockley
// problems: one launch; each workgroup takes an instance from a work queue (largest
// first) and runs the whole Picard loop of ElectricalFem2DSolver::compute
// (solvers/electrical/shockley/electr2d.cpp:363-440) for it on the device:
//   junction update (electr2d.cpp:243-264, law beta.hpp:43-46)
//   -> line-preconditioned CG on the eliminated system (warm start)
//   -> element currents and the loop error (electr2d.cpp:399-423).
// Inside a workgroup: node-centric operator apply (grid-stride, all threads),
// Thomas line solves along axis 1 (one thread per column), LDS reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_kernels.hpp"

namespace pf {

struct BatchInst {
    int n0, n1;                  // nodes along r (axis 0) and z (axis 1)
    long long node_off, elem_off;
    int h0_off, h1_off;          // offsets of this instance's element sizes in h0[] (and rmid[]) / h1[]
    int jlo, jhi;                // junction element rows [jlo, jhi); jlo == jhi: none
    double jheight, beta, js;
    int cyl;
    int nlev;     // multigrid levels below the instance grid (coarse levels 1..nlev)
    int lev_off;  // first of them in the level table
    int pad;
};

// One coarse multigrid level of an instance (axis 0 coarsened, every row of axis 1 kept);
// offsets into the batch's level pools.  The transfer tables describe the step from the
// next finer level: cx[I] = fine index of coarse node I; fine node i lies in coarse
// interval [mx[i], mx[i] + 1] with right weight wx[i].
struct BatchLevel {
    int n0;
    int cx_off, mx_off;  // into the int tables (cx) and the int / double tables (mx, wx)
    int h0_off;          // element widths and mid radii of this level
    long long node_off;  // level node pool (b, x, r, line factors)
    long long elem_off;  // level element pool (conductivities)
    long long mask_off;  // bit index of node 0 in the level Dirichlet mask
};

// Pools of the coarse levels of all instances (see BatchLevel).
struct BatchMg {
    const BatchLevel* lev;
    const int* cx;
    const int* mx;
    const double* wx;
    const double* h0;
    const double* rm;
    const uint64_t* mask;
    double *kh, *kv, *b, *x, *r, *la, *lw, *lc;
};

struct BatchStats {
    int loops;      // Picard loops done
    int iters;      // CG iterations summed over loops
    int status;     // 0 ok, 1 CG maxit, 2 Picard maxloops, <0 breakdown
    int restarts;   // CG restarts from a recomputed true residual, summed over loops
    double err;     // last loop error (%)
    double maxj;    // max |j| at the junction (kA/cm^2)
    double rel;     // last loop: final true residual over its reference residual
    double current; // total current through the junction middle row (mA), cylindrical pi sum j (r2^2 - r1^2) 0.01
    double bwd;     // last loop: componentwise backward error max_i |r_i| / (|A0| |V|)_i of its solution
};

struct BatchParams {
    int maxloops, maxit;
    double maxerr, rtol;
    int stable;
    int precond;  // 2 line, 3 multigrid V-cycle
    double omega; // multigrid smoother damping
    int coarse_sweeps;
    int loop_budget;  // Picard loops one launch runs per instance (0: the whole loop)
};

// What an instance carries from one launch to the next: the denominator of its CG
// stopping test (the residual norm of the start field) and whether its Picard loop
// has ended.  The loop count, status and iterations live in BatchStats; potentials,
// conductivities and currents stay in their device arrays.
struct BatchResume {
    double bb;
    int done;
    int pad;
};

__device__ __forceinline__ double wg_sum(double v, double* red) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = (blockDim.x + 63) >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.;
    for (int w = 0; w < nw; ++w) s += red[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double wg_max(double v, double* red) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = (blockDim.x + 63) >> 6;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = -__builtin_huge_val();
    for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

// Error-free transformations (Knuth's TwoSum; TwoProduct with a fused multiply-add,
// Ogita, Rump & Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26, 2005).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bv = s - a;
    e = (a - (s - bv)) + (b - bv);
}

// One level of an instance as the workgroup sees it (level 0: the instance grid).
struct LView {
    int n0, n1, cyl;
    const double* h0;  // element widths along axis 0 (and mid radii rm)
    const double* h1;  // element heights along axis 1
    const double* rm;
    double* kh;
    double* kv;
    const uint64_t* mask;
    long long mbase;  // bit index of node 0 in mask
    double *la, *lw, *lc;
};

__device__ __forceinline__ bool lbit(const LView& V, long long node) { return dbit(V.mask, V.mbase + node); }

// Operator rows of a level marched along axis 1.  A work item is one column i and a range
// of rows (columns times row ranges cover at least the block); it carries the 3 x 3 node
// window around its node and the element coefficients of the two element rows there in
// registers, so a node costs three potential loads, four conductivity loads and three
// mask bits instead of the ~30 loads of a node-centric row, and the loads of the next row
// are issued before the current row is formed.  Element matrix of therm2d.cpp:225-228 /
// electr2d.cpp:296-304 (x r_mid for cylindrical geometry), summed element by element.
// KIND 0: y = M A0 M x + (I - M) x; returns the sum of x y over the rows formed;
// KIND 1: y = b + M (0 - A0 x) (b null: 0); returns 0;
// KIND 2: y = M (0 - A0 x) with the products and their sum carried in double-double (Dot2):
//   the residual of a nearly converged potential is then accurate to about the rounding of
//   the result, not to eps |A0| |x|, so CG restarted from it can reduce the true residual
//   below the level its recursive residual drifts to; returns the sum of y^2, and bw the max
//   over the rows of |y_i| / (|A0| |x|)_i (no FP64 potential has a residual much below eps
//   times that row).
template <int BS, int KIND>
__device__ inline double march_l(const LView& V, const double* __restrict__ x, const double* __restrict__ b,
                                 double* __restrict__ y, double& bw) {
    const int n0 = V.n0, n1 = V.n1;
    const int G = n0 >= BS ? 1 : BS / n0;  // row ranges per column
    const int R = (n1 + G - 1) / G;
    double acc = 0.;
    bw = 0.;
    for (int w = threadIdx.x; w < G * n0; w += BS) {
        const int g = w / n0, i = w - g * n0;
        const int k0 = g * R, k1 = min(n1, k0 + R);
        if (k0 >= k1) continue;
        const bool eL = i >= 1, eR = i < n0 - 1;
        const int im = max(i - 1, 0), ip = min(i + 1, n0 - 1);
        const int eiL = max(i - 1, 0), eiR = min(i, n0 - 2);
        const double wL = V.h0[eiL], wR = V.h0[eiR];
        const double rL = V.cyl ? V.rm[eiL] : 1., rR = V.cyl ? V.rm[eiR] : 1.;
        // node row kk (clamped) into xr / mr
        auto load_row = [&](int kk, double (&xr)[3], bool (&mr)[3]) {
            const int kc = min(max(kk, 0), n1 - 1);
            const long long base = (long long)kc * n0;
            xr[0] = x[base + im];
            xr[1] = x[base + i];
            xr[2] = x[base + ip];
            mr[0] = lbit(V, base + im);
            mr[1] = lbit(V, base + i);
            mr[2] = lbit(V, base + ip);
        };
        // element row ek (clamped): raw conductivities and the height
        auto load_elems = [&](int ek, double (&kr)[4], double& hr) {
            const int ec = min(max(ek, 0), n1 - 2);
            const long long eb = (long long)ec * (n0 - 1);
            kr[0] = V.kh[eb + eiL];
            kr[1] = V.kv[eb + eiL];
            kr[2] = V.kh[eb + eiR];
            kr[3] = V.kv[eb + eiR];
            hr = V.h1[ec];
        };
        double xw[3][3], xn[3], kn[4], hn, ke[2][4], he[2];
        bool mw[3][3], mn[3];
        load_row(k0 - 1, xw[0], mw[0]);
        load_row(k0, xw[1], mw[1]);
        load_elems(k0 - 1, ke[0], he[0]);
        load_row(k0 + 1, xn, mn);
        load_elems(k0, kn, hn);
        for (int k = k0; k < k1; ++k) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xw[2][d] = xn[d];
                mw[2][d] = mn[d];
            }
#pragma unroll
            for (int d = 0; d < 4; ++d) ke[1][d] = kn[d];
            he[1] = hn;
            load_row(k + 2, xn, mn);
            load_elems(k + 1, kn, hn);
            const long long node = (long long)k * n0 + i;
            const bool di = mw[1][1];
            double out;
            if (KIND == 2) {
                double hi = 0., lo = 0., mag = 0.;
                if (!di) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int cx = c & 1, cy = c >> 1;
                        const bool ok = (cx ? eR : eL) && (cy ? (k < n1 - 1) : (k >= 1));
                        if (!ok) continue;
                        const double wv = cx ? wR : wL, rv = cx ? rR : rL, h = he[cy];
                        const double kx = rv * ke[cy][2 * cx] * h / wv, ky = rv * ke[cy][2 * cx + 1] * wv / h;
                        const int ax = cx ? 0 : 1, ay = cy ? 0 : 1;
#pragma unroll
                        for (int bb = 0; bb < 4; ++bb) {
                            const int bx = bb & 1, by = bb >> 1;
                            const double Sx = (ax == bx) ? 1. : -1., Sy = (ay == by) ? 1. : -1.;
                            const double Mx = (ax == bx) ? (1. / 3.) : (1. / 6.), My = (ay == by) ? (1. / 3.) : (1. / 6.);
                            const double a = kx * Sx * My + ky * Mx * Sy, xv = xw[cy + by][cx + bx];
                            const double pr = a * xv, pe = fma(a, xv, -pr);
                            double se;
                            two_sum(hi, pr, hi, se);
                            lo += pe + se;
                            mag += fabs(pr);
                        }
                    }
                }
                out = di ? 0. : -(hi + lo);
                acc += out * out;
                if (mag > 0.) bw = fmax(bw, fabs(out) / mag);
            } else {
                double s = 0.;
                if (!di) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int cx = c & 1, cy = c >> 1;
                        const bool ok = (cx ? eR : eL) && (cy ? (k < n1 - 1) : (k >= 1));
                        if (!ok) continue;
                        const double wv = cx ? wR : wL, rv = cx ? rR : rL, h = he[cy];
                        const double kx = rv * ke[cy][2 * cx] * h / wv, ky = rv * ke[cy][2 * cx + 1] * wv / h;
                        const int ax = cx ? 0 : 1, ay = cy ? 0 : 1;
#pragma unroll
                        for (int bb = 0; bb < 4; ++bb) {
                            const int bx = bb & 1, by = bb >> 1;
                            if (KIND == 0 && mw[cy + by][cx + bx]) continue;
                            const double Sx = (ax == bx) ? 1. : -1., Sy = (ay == by) ? 1. : -1.;
                            const double Mx = (ax == bx) ? (1. / 3.) : (1. / 6.), My = (ay == by) ? (1. / 3.) : (1. / 6.);
                            s += (kx * Sx * My + ky * Mx * Sy) * xw[cy + by][cx + bx];
                        }
                    }
                }
                if (KIND == 0) {
                    out = di ? xw[1][1] : s;
                    acc += xw[1][1] * out;
                } else {
                    out = (b ? b[node] : 0.) + (di ? 0. : -s);
                }
            }
            y[node] = out;
            // shift the window one row up
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xw[0][d] = xw[1][d];
                mw[0][d] = mw[1][d];
                xw[1][d] = xw[2][d];
                mw[1][d] = mw[2][d];
            }
#pragma unroll
            for (int d = 0; d < 4; ++d) ke[0][d] = ke[1][d];
            he[0] = he[1];
        }
    }
    return acc;
}

// Thomas factors of the tridiagonal part of the level's matrix along axis 1 (per column)
template <int BS>
__device__ inline void line_factors_l(const LView& V) {
    const int tid = threadIdx.x;
    for (int i = tid; i < V.n0; i += BS) {
        double cprev = 0., cpprev = 0.;
        for (int k = 0; k < V.n1; ++k) {
            const long long node = (long long)k * V.n0 + i;
            double b = 0., c = 0.;
            for (int cc = 0; cc < 4; ++cc) {
                const int ei = i - 1 + (cc & 1), ek = k - 1 + (cc >> 1);
                if (ei < 0 || ek < 0 || ei >= V.n0 - 1 || ek >= V.n1 - 1) continue;
                const long long e = (long long)ek * (V.n0 - 1) + ei;
                const double w = V.h0[ei], h = V.h1[ek];
                const double rr = V.cyl ? V.rm[ei] : 1.;
                const double kx1 = rr * V.kh[e] * h / w, ky1 = rr * V.kv[e] * w / h;
                b += (kx1 + ky1) / 3.;
                if (ek == k) c += (kx1 - 2. * ky1) / 6.;
            }
            const bool dk = lbit(V, node);
            if (dk) b = 1.;
            if (dk || k == V.n1 - 1 || lbit(V, node + V.n0)) c = 0.;
            const double a = cprev, w = 1. / (b - a * cpprev), cp = c * w;
            V.la[node] = a;
            V.lw[node] = w;
            V.lc[node] = cp;
            cprev = c;
            cpprev = cp;
        }
    }
    __syncthreads();
}

// xout = xin + om M^-1 rhs (xin null: from zero; xin may be xout; rhs may be tmp),
// the line solve per column with the forward sweep kept in tmp
template <int BS>
__device__ inline void line_sweep_l(const LView& V, const double* rhs, double* tmp, const double* xin, double* xout,
                                    double om) {
    // The loads of a chunk of U planes do not depend on the recurrence: they are all issued
    // before its U steps, and the next chunk's before the current chunk's steps (two
    // register buffers), so the sweep waits for memory about once per two chunks.  A chunk
    // is read before any of its nodes is written, and the chunk read ahead lies beyond the
    // rows being written (rhs may be tmp, xin may be xout).  Reads past either end are
    // clamped to the last row and unused.
    constexpr int U = 8;
    const int tid = threadIdx.x;
    const int n1 = V.n1;
    for (int i = tid; i < V.n0; i += BS) {
        double d = 0.;
        double ra[U], aa[U], wa[U], rb[U], ab[U], wb[U];
        auto ldf = [&](int k0, double (&rv)[U], double (&av)[U], double (&wv)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(k0 + u, n1 - 1);
                const long long node = (long long)k * V.n0 + i;
                rv[u] = rhs[node];
                av[u] = V.la[node];
                wv[u] = V.lw[node];
            }
        };
        auto stf = [&](int k0, const double (&rv)[U], const double (&av)[U], const double (&wv)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (k0 + u < n1) {
                    d = (rv[u] - av[u] * d) * wv[u];
                    tmp[(long long)(k0 + u) * V.n0 + i] = d;
                }
            }
        };
        ldf(0, ra, aa, wa);
        for (int k0 = 0; k0 < n1; k0 += 2 * U) {
            ldf(k0 + U, rb, ab, wb);
            stf(k0, ra, aa, wa);
            ldf(k0 + 2 * U, ra, aa, wa);
            stf(k0 + U, rb, ab, wb);
        }
        double zn = 0.;
        double ta[U], ca[U], xa[U], tb[U], cb[U], xb[U];
        auto ldb = [&](int k0, double (&tv)[U], double (&cv)[U], double (&xv)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = max(k0 - u, 0);
                const long long node = (long long)k * V.n0 + i;
                tv[u] = tmp[node];
                cv[u] = V.lc[node];
                xv[u] = xin ? xin[node] : 0.;
            }
        };
        auto stb = [&](int k0, const double (&tv)[U], const double (&cv)[U], const double (&xv)[U]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (k0 - u >= 0) {
                    zn = tv[u] - cv[u] * zn;
                    xout[(long long)(k0 - u) * V.n0 + i] = xv[u] + om * zn;
                }
            }
        };
        ldb(n1 - 1, ta, ca, xa);
        for (int k0 = n1 - 1; k0 >= 0; k0 -= 2 * U) {
            ldb(k0 - U, tb, cb, xb);
            stb(k0, ta, ca, xa);
            ldb(k0 - 2 * U, ta, ca, xa);
            stb(k0 - U, tb, cb, xb);
        }
    }
    __syncthreads();
}

// r = b - A x on a level (zero at Dirichlet nodes, where b is zero)
template <int BS>
__device__ inline void residual_l(const LView& V, const double* b, const double* x, double* r) {
    double bw;
    march_l<BS, 1>(V, x, b, r, bw);
    __syncthreads();
}

// b_c = P^T r_f (coarse Dirichlet nodes: 0) and x_f += P x_c (fine Dirichlet nodes skipped)
// The transfers run over work items of one column and a range of rows (as march_l); the
// column's tables are read once, and the rows go in groups of 4 whose loads are all issued
// before any of their stores (independent rows: memory waits once per group, not per row).
template <int BS>
__device__ inline void restrict_l(const LView& F, const LView& Cv, const int* cx, const int* mx, const double* wx,
                                  const double* __restrict__ rf, double* __restrict__ bc) {
    const int n0 = Cv.n0, n1 = Cv.n1;
    const int G = n0 >= BS ? 1 : BS / n0;
    const int R = (n1 + G - 1) / G;
    for (int w = threadIdx.x; w < G * n0; w += BS) {
        const int g = w / n0, I = w - g * n0;
        const int k0 = g * R, k1 = min(n1, k0 + R);
        const int ilo = I > 0 ? cx[I - 1] + 1 : 0, ihi = I < n0 - 1 ? cx[I + 1] - 1 : cx[I];
        // up to 3 fine columns (absent ones: weight 0 at a valid column)
        int ii[3];
        double wt[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int f = ilo + a;
            const bool ok = f <= ihi;
            ii[a] = ok ? f : ilo;
            wt[a] = ok ? tweight(mx[ii[a]], wx[ii[a]], I) : 0.;
        }
        constexpr int U = 4;
        for (int kb = k0; kb < k1; kb += U) {
            double v[U][3];
            bool m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(kb + u, k1 - 1);
                const double* row = rf + (long long)k * F.n0;
                v[u][0] = row[ii[0]];
                v[u][1] = row[ii[1]];
                v[u][2] = row[ii[2]];
                m[u] = lbit(Cv, (long long)k * n0 + I);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kb + u < k1)
                    bc[(long long)(kb + u) * n0 + I] = m[u] ? 0. : wt[0] * v[u][0] + wt[1] * v[u][1] + wt[2] * v[u][2];
            }
        }
    }
    __syncthreads();
}

template <int BS>
__device__ inline void prolong_l(const LView& F, const LView& Cv, const int* mx, const double* wx,
                                 const double* __restrict__ xc, double* __restrict__ xf) {
    const int n0 = F.n0, n1 = F.n1;
    const int G = n0 >= BS ? 1 : BS / n0;
    const int R = (n1 + G - 1) / G;
    for (int w = threadIdx.x; w < G * n0; w += BS) {
        const int g = w / n0, i = w - g * n0;
        const int k0 = g * R, k1 = min(n1, k0 + R);
        const int I = mx[i], I1 = min(I + 1, Cv.n0 - 1);
        const double a = wx[i];
        constexpr int U = 4;
        for (int kb = k0; kb < k1; kb += U) {
            double xo[U], c0[U], c1[U];
            bool m[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(kb + u, k1 - 1);
                xo[u] = xf[(long long)k * n0 + i];
                c0[u] = xc[(long long)k * Cv.n0 + I];
                c1[u] = xc[(long long)k * Cv.n0 + I1];
                m[u] = lbit(F, (long long)k * n0 + i);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kb + u < k1 && !m[u]) xf[(long long)(kb + u) * n0 + i] = xo[u] + ((1. - a) * c0[u] + a * c1[u]);
            }
        }
    }
    __syncthreads();
}

// coarse conductivities of a level from the next finer one: width (x mid radius) means
template <int BS>
__device__ inline void coarsen_k_l(const LView& F, const LView& Cv, const int* cx) {
    const long long Ec = (long long)(Cv.n0 - 1) * (Cv.n1 - 1);
    for (long long e = threadIdx.x; e < Ec; e += BS) {
        const int I = (int)(e % (Cv.n0 - 1));
        const long long k = e / (Cv.n0 - 1);
        double sa = 0., sh = 0., sv = 0.;
        for (int ex = cx[I]; ex < cx[I + 1]; ++ex) {
            const double a = F.h0[ex] * (F.cyl ? F.rm[ex] : 1.);
            const long long f = k * (F.n0 - 1) + ex;
            sa += a;
            sh += a * F.kh[f];
            sv += a * F.kv[f];
        }
        Cv.kh[e] = sh / sa;
        Cv.kv[e] = sv / sa;
    }
    __syncthreads();
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch_shockley2d(const BatchInst* __restrict__ insts, const int* __restrict__ order,
                                                         int ninst, unsigned int* __restrict__ queue, BatchParams prm,
                                                         const double* __restrict__ h0, const double* __restrict__ h1,
                                                         const double* __restrict__ rm, double* __restrict__ kh,
                                                         double* __restrict__ kv, const uint64_t* __restrict__ dmask,
                                                         double* __restrict__ X, double* __restrict__ R,
                                                         double* __restrict__ P, double* __restrict__ Q,
                                                         double* __restrict__ Z, double* __restrict__ LA,
                                                         double* __restrict__ LW, double* __restrict__ LC,
                                                         double* __restrict__ CUR, BatchStats* __restrict__ stats,
                                                         BatchResume* __restrict__ resume, BatchMg mg) {
    __shared__ double red[16];
    __shared__ int s_q;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_q = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int qi = s_q;
        __syncthreads();
        if (qi >= ninst) break;
        const int id = order[qi];
        BatchResume RS = resume[id];
        if (RS.done) continue;  // uniform across the workgroup: every thread read the same entry
        const BatchInst I = insts[id];
        const int n0 = I.n0, n1 = I.n1;
        const long long N = (long long)n0 * n1, E = (long long)(n0 - 1) * (n1 - 1);
        double* x = X + I.node_off;
        double* r = R + I.node_off;
        double* p = P + I.node_off;
        double* q = Q + I.node_off;
        double* z = Z + I.node_off;
        double* la = LA + I.node_off;
        double* lw = LW + I.node_off;
        double* lc = LC + I.node_off;
        double* cur = CUR + 2 * I.elem_off;
        double* sh = kh + I.elem_off;
        double* sv = kv + I.elem_off;
        const bool has_j = I.jhi > I.jlo;
        BatchStats S = stats[id];  // zero before the first launch, then the state the last launch left
        // multigrid: level views (0 = this grid) and their b / x / scratch vectors; on the
        // instance grid b is the CG residual, x the preconditioned residual z, the
        // scratch q (free between the update and the next operator apply)
        const bool mgon = prm.precond == 3;
        const LView V0{n0, n1, I.cyl, h0 + I.h0_off, h1 + I.h1_off, rm + I.h0_off, sh, sv, dmask, I.node_off, la, lw, lc};
        auto lview = [&](int l) -> LView {
            if (l == 0) return V0;
            const BatchLevel& B = mg.lev[I.lev_off + l - 1];
            return LView{B.n0, n1, I.cyl, mg.h0 + B.h0_off, h1 + I.h1_off, mg.rm + B.h0_off, mg.kh + B.elem_off,
                         mg.kv + B.elem_off, mg.mask, B.mask_off, mg.la + B.node_off, mg.lw + B.node_off,
                         mg.lc + B.node_off};
        };
        auto lb = [&](int l) { return l ? mg.b + mg.lev[I.lev_off + l - 1].node_off : r; };
        auto lx = [&](int l) { return l ? mg.x + mg.lev[I.lev_off + l - 1].node_off : z; };
        auto lt = [&](int l) { return l ? mg.r + mg.lev[I.lev_off + l - 1].node_off : q; };
        // z = V r: damped line sweep before and after each coarse correction, coarse sweeps
        // on the coarsest level (the V-cycle of the single-instance solver)
        auto vcycle = [&]() {
            const int L = I.nlev + 1;
            const double om = prm.omega;
            for (int l = 0; l < L; ++l) {
                const LView V = lview(l);
                double *bl = lb(l), *xl = lx(l), *tl = lt(l);
                line_sweep_l<BS>(V, bl, tl, nullptr, xl, om);
                if (l == L - 1) {
                    for (int sw = 1; sw < prm.coarse_sweeps; ++sw) {
                        residual_l<BS>(V, bl, xl, tl);
                        line_sweep_l<BS>(V, tl, tl, xl, xl, om);
                    }
                } else {
                    residual_l<BS>(V, bl, xl, tl);
                    const BatchLevel& B = mg.lev[I.lev_off + l];
                    restrict_l<BS>(V, lview(l + 1), mg.cx + B.cx_off, mg.mx + B.mx_off, mg.wx + B.mx_off, tl, lb(l + 1));
                }
            }
            for (int l = L - 2; l >= 0; --l) {
                const LView V = lview(l);
                double *bl = lb(l), *xl = lx(l), *tl = lt(l);
                const BatchLevel& B = mg.lev[I.lev_off + l];
                prolong_l<BS>(V, lview(l + 1), mg.mx + B.mx_off, mg.wx + B.mx_off, lx(l + 1), xl);
                residual_l<BS>(V, bl, xl, tl);
                line_sweep_l<BS>(V, tl, tl, xl, xl, om);
            }
        };
        // denominator: residual of the start field (Dirichlet values, zero elsewhere),
        // computed by the launch that starts the instance
        double bb = RS.bb;
        if (S.loops == 0) {
            double bw0;
            march_l<BS, 1>(V0, x, nullptr, r, bw0);
            __syncthreads();
            bb = 0.;
            for (long long n = tid; n < N; n += BS) bb += r[n] * r[n];
            bb = wg_sum(bb, red);
        }
        const double bbs = bb > 0. ? bb : 1.;
        double err = S.err, mcur = S.maxj;
        bool finished = false;
        const int loop_end = prm.loop_budget > 0 ? min(prm.maxloops, S.loops + prm.loop_budget) : prm.maxloops;
        for (int loop = S.loops; loop < loop_end; ++loop) {
            // junction update from the current potential (not before the first solve)
            if (loop > 0 && has_j) {
                const long long nje = (long long)(I.jhi - I.jlo) * (n0 - 1);
                for (long long t = tid; t < nje; t += BS) {
                    const int ei = t % (n0 - 1), ek = I.jlo + t / (n0 - 1);
                    const long long e = (long long)ek * (n0 - 1) + ei;
                    const double U = 0.5 * (x[(long long)I.jhi * n0 + ei] - x[(long long)I.jlo * n0 + ei] +
                                            x[(long long)I.jhi * n0 + ei + 1] - x[(long long)I.jlo * n0 + ei + 1]);
                    const double jy = fabs(0.1 * sv[e] * U / I.jheight);
                    double cond = 10. * jy * I.jheight * I.beta / log(1e7 * jy / I.js + 1.);
                    if (prm.stable) cond = 0.5 * (sv[e] + cond);
                    if (isnan(cond) || fabs(cond) < 1e-16) cond = 1e-16;
                    sv[e] = cond;
                    sh[e] = 0.;
                }
                __syncthreads();
            }
            // line factors along axis 1 (tridiagonal part of A per column); with the
            // multigrid preconditioner also the coarse conductivities and line factors
            line_factors_l<BS>(V0);
            if (mgon)
                for (int l = 1; l <= I.nlev; ++l) {
                    const LView F = lview(l - 1), Cv = lview(l);
                    coarsen_k_l<BS>(F, Cv, mg.cx + mg.lev[I.lev_off + l - 1].cx_off);
                    line_factors_l<BS>(Cv);
                }
            // CG from the current potential
            // true residual (compensated) and the componentwise backward error of x
            double bwd = 0.;
            auto true_residual = [&]() {
                double bw = 0.;
                const double acc = march_l<BS, 2>(V0, x, nullptr, r, bw);
                bwd = wg_max(bw, red);
                return wg_sum(acc, red);
            };
            double rz = 0., rr = true_residual();
            auto precond = [&]() {
                double acc = 0.;
                if (mgon) {
                    vcycle();
                    for (long long n = tid; n < N; n += BS) acc += r[n] * z[n];
                    return wg_sum(acc, red);
                }
                for (int i = tid; i < n0; i += BS) {
                    double d = 0.;
                    for (int k = 0; k < n1; ++k) {
                        const long long node = (long long)k * n0 + i;
                        d = (r[node] - la[node] * d) * lw[node];
                        z[node] = d;
                    }
                    double zn = 0.;
                    for (int k = n1 - 1; k >= 0; --k) {
                        const long long node = (long long)k * n0 + i;
                        zn = z[node] - lc[node] * zn;
                        z[node] = zn;
                        acc += r[node] * zn;
                    }
                }
                __syncthreads();
                return wg_sum(acc, red);
            };
            // each loop solves its own system: its start residual (the previous loop's
            // potential against the updated junction) reduced by rtol, as a direct solve
            // would per loop, but never below rtol * 1e-4 of the start field's residual
            const double stop2 = prm.rtol * prm.rtol * fmax(rr, 1e-8 * bbs);
            int it = 0, cst = 0;  // this loop's CG status (0, 1 maxit, -6 / -7 breakdown)
            // CG runs from the true residual; when the recursive residual meets the
            // tolerance, the true one is recomputed (compensated) and CG restarts once from
            // it if it has drifted above the tolerance.  On junction systems whose
            // conductivities span ~1e14 no FP64 Krylov run reaches a late loop's tolerance;
            // the potential is then as accurate as FP64 CG makes it (measured: within ~1e-8
            // relative of a band-Cholesky solve of the same system), which `bwd` reports.
            double rr_run = rr;
            for (int run = 0; run < 2 && rr > stop2 && cst == 0; ++run) {
                if (run > 0) {
                    if (rr > 0.25 * rr_run) break;
                    ++S.restarts;  // summed over loops
                    rr_run = rr;
                }
                rz = precond();
                for (long long n = tid; n < N; n += BS) p[n] = z[n];
                __syncthreads();
                for (;;) {
                    double bwq;
                    double pq = march_l<BS, 0>(V0, p, nullptr, q, bwq);
                    pq = wg_sum(pq, red);
                    if (!(pq > 0.)) {
                        cst = -6;
                        break;
                    }
                    const double alpha = rz / pq;
                    rr = 0.;
                    for (long long n = tid; n < N; n += BS) {
                        x[n] += alpha * p[n];
                        const double v = r[n] - alpha * q[n];
                        r[n] = v;
                        rr += v * v;
                    }
                    rr = wg_sum(rr, red);
                    ++it;
                    if (rr <= stop2) break;
                    if (it >= prm.maxit) {
                        cst = 1;
                        break;
                    }
                    const double rz_new = precond();
                    if (!(rz_new > 0.)) {
                        cst = -7;
                        break;
                    }
                    const double beta = rz_new / rz;
                    rz = rz_new;
                    for (long long n = tid; n < N; n += BS) p[n] = z[n] + beta * p[n];
                    __syncthreads();
                }
                if (cst == 0) rr = true_residual();
            }
            if (cst != 0) S.status = cst;
            S.iters += it;
            S.rel = sqrt(rr / stop2) * prm.rtol;  // final residual over the loop's reference residual
            S.bwd = bwd;
            // currents and the loop error
            double dmax = 0., jmax = 0.;
            for (long long e = tid; e < E; e += BS) {
                const int ei = e % (n0 - 1), ek = e / (n0 - 1);
                const double w = h0[I.h0_off + ei], h = h1[I.h1_off + ek];
                const long long ll = (long long)ek * n0 + ei;
                const double vll = x[ll], vlu = x[ll + 1], vul = x[ll + n0], vuu = x[ll + n0 + 1];
                const double jx = sh[e] * (-0.05 * (-vll + vlu - vul + vuu) / w);
                const double jy = sv[e] * (-0.05 * (-vll - vlu + vul + vuu) / h);
                const double dx = jx - cur[2 * e], dy = jy - cur[2 * e + 1];
                cur[2 * e] = jx;
                cur[2 * e + 1] = jy;
                dmax = fmax(dmax, dx * dx + dy * dy);
                if (!has_j || (ek >= I.jlo && ek < I.jhi)) jmax = fmax(jmax, jx * jx + jy * jy);
            }
            dmax = wg_max(dmax, red);
            jmax = wg_max(jmax, red);
            mcur = sqrt(jmax);
            err = 100. * sqrt(dmax) / fmax(mcur, 100e-7);
            S.loops = loop + 1;
            if (S.status < 0 || (S.status == 0 && err <= prm.maxerr)) {
                finished = true;
                break;
            }
            if (loop + 1 == prm.maxloops) {
                S.status = 2;
                finished = true;
            }
        }
        // total current through the middle junction row (integrateCurrent, electr2d.cpp:466-496)
        double tot = 0.;
        if (has_j) {
            const int level = (I.jlo + I.jhi) / 2;
            for (int ei = tid; ei < n0 - 1; ei += BS) {
                const long long e = (long long)level * (n0 - 1) + ei;
                const double j1 = cur[2 * e + 1];
                if (I.cyl) {
                    double r0 = 0., r1 = 0.;
                    // element radii from cumulative sizes: rmid +- w/2
                    const double w = h0[I.h0_off + ei];
                    r0 = rm[I.h0_off + ei] - 0.5 * w;
                    r1 = rm[I.h0_off + ei] + 0.5 * w;
                    tot += j1 * (r1 * r1 - r0 * r0);
                } else {
                    tot += j1 * h0[I.h0_off + ei];
                }
            }
        }
        tot = wg_sum(tot, red);
        if (tid == 0) {
            S.err = err;
            S.maxj = mcur;
            S.current = I.cyl ? tot * 3.14159265358979323846 * 0.01 : tot * 0.01;
            stats[id] = S;
            RS.bb = bb;
            RS.done = finished ? 1 : 0;
            resume[id] = RS;
            if (finished) atomicAdd(queue + 1, 1u);  // instances finished, read by the host between launches
        }
        __syncthreads();
    }
}

__global__ void k_reset_counter(unsigned int* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// Start of a batch solve: every instance not started, no instance finished.
__global__ void k_batch_reset(int ninst, BatchStats* __restrict__ stats, BatchResume* __restrict__ resume,
                              unsigned int* __restrict__ queue) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) {
        stats[i] = BatchStats{};
        resume[i] = BatchResume{};
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = 0u;
        queue[1] = 0u;
    }
}

}  // namespace pf
