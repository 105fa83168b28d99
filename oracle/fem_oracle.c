/*
 * CPU oracle for the FEM thermal / electrical steady-state path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against, and the CPU baseline timed beside it.  Nothing in the
 * product (plask_amd/) links, loads or calls it.
 *
 * It restates, in plain C and FP64, the reference algorithms:
 *   - Q1 element stiffness + load (solvers/thermal/static/therm2d.cpp:216-228,
 *     :342-354 (cylindrical x r), therm3d.cpp:209-240;
 *     solvers/electrical/shockley/electr2d.cpp:287-316, electr3d.cpp:310-341)
 *   - LAPACK-'L' symmetric band storage (plask/common/fem/cholesky_matrix.hpp:58-88)
 *     and NSPCG symmetric-diagonal storage (plask/common/fem/iterative_matrix.hpp:347-433)
 *   - Dirichlet elimination (plask/common/fem/matrix.hpp:132-145,
 *     iterative_matrix.hpp:462-485)
 *   - band Cholesky factor/solve (the dpbtrf/dpbtrs pair called at
 *     cholesky_matrix.hpp:95,107), textbook right-looking form
 *   - preconditioned CG as iterated by NSPCG itcg (extlib/nspcg/nspcg.f:9217-9330)
 *     with Jacobi (pjac, nspcg.f:25615) or IC(0) preconditioning, stopped on the
 *     true relative residual ||B - A X|| / ||B||
 *   - Shockley junction law (solvers/electrical/shockley/beta.hpp:43-46) and the
 *     element current / Joule heat post-processing (electr2d.cpp:399-420,442-464;
 *     electr3d.cpp:389-478)
 *
 * Element arrays are in natural element order e = i0 + (n0-1)*(i1 + (n1-1)*i2).
 * Node numbering is given by strides: node(i0,i1,i2) = i0*s0 + i1*s1 + i2*s2,
 * which covers every iteration order of the reference rectilinear meshes
 * (plask/mesh/rectilinear3d.cpp:20-23, rectangular2d.hpp:296-297).
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    int dim;             /* 2 or 3 */
    int cyl;             /* 2D only: axis 0 is the radius, element matrix x r_mid */
    int64_t n[3];        /* nodes per axis */
    const double* ax[3]; /* node coordinates per axis (um) */
    int64_t s[3];        /* node strides */
} orc_grid;

#define ST_BAND 0 /* LAPACK 'L' band: A(r>=c) at ld*c + r */
#define ST_DIA 1  /* NSPCG symmetric diagonal storage: A(c+off_d, c) at c + rank*d */

typedef struct {
    int kind;
    int64_t rank;
    int64_t ld;             /* band: leading dimension - 1 */
    int64_t kd;             /* band: half bandwidth */
    int ndiag;              /* dia: number of stored diagonals */
    const int64_t* offsets; /* dia: offsets[0] == 0 */
    double* data;
} orc_mat;

static double* mat_ref(orc_mat* A, int64_t r, int64_t c) {
    if (A->kind == ST_BAND) {
        if (r < c) { int64_t t = r; r = c; c = t; }
        return A->data + A->ld * c + r;
    } else {
        if (r == c) return A->data + r;
        if (r < c) { int64_t t = r; r = c; c = t; }
        int64_t off = r - c;
        for (int d = 1; d < A->ndiag; ++d)
            if (A->offsets[d] == off) return A->data + c + A->rank * d;
        return NULL;
    }
}

int64_t orc_elem_count(const orc_grid* g) {
    int64_t e = (g->n[0] - 1) * (g->n[1] - 1);
    if (g->dim == 3) e *= (g->n[2] - 1);
    return e;
}

int64_t orc_node_count(const orc_grid* g) {
    int64_t n = g->n[0] * g->n[1];
    if (g->dim == 3) n *= g->n[2];
    return n;
}

/* 2D Q1 element matrix: therm2d.cpp:216-228 / electr2d.cpp:290-301.
 * Local order 1=(lo,lo) 2=(up,lo) 3=(up,up) 4=(lo,up) (therm2d.cpp:196-199). */
static void elem_K2(double kx, double ky, double w, double h, double K[4][4]) {
    kx = kx * h / w;
    ky = ky * w / h;
    double k11 = (kx + ky) / 3., k21 = (-2. * kx + ky) / 6., k31 = -(kx + ky) / 6., k32 = (kx - 2. * ky) / 6.;
    /* k44=k33=k22=k11; k43=k21; k42=k31; k41=k32 */
    double M[4][4] = {{k11, k21, k31, k32}, {k21, k11, k32, k31}, {k31, k32, k11, k21}, {k32, k31, k21, k11}};
    memcpy(K, M, sizeof(M));
}

/* 3D Q1 element matrix: therm3d.cpp:218-237 (kx, ky, kz already in W/um/K).
 * Local node bit0 = x, bit1 = y, bit2 = z (therm3d.cpp:189-197). */
static void elem_K3(double kx, double ky, double kz, double dx, double dy, double dz, double K[8][8]) {
    kx = kx / dx * dy * dz;
    ky = ky * dx / dy * dz;
    kz = kz * dx * dy / dz;
    double c[8];
    c[0] = (kx + ky + kz) / 9.;
    c[1] = (-2. * kx + ky + kz) / 18.;
    c[2] = (kx - 2. * ky + kz) / 18.;
    c[4] = (kx + ky - 2. * kz) / 18.;
    c[6] = (kx - 2. * ky - 2. * kz) / 36.;
    c[5] = (-2. * kx + ky - 2. * kz) / 36.;
    c[3] = (-2. * kx - 2. * ky + kz) / 36.;
    c[7] = -(kx + ky + kz) / 36.;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) K[i][j] = c[i ^ j];
}

static int64_t eidx(const orc_grid* g, int64_t i0, int64_t i1, int64_t i2) {
    return i0 + (g->n[0] - 1) * (i1 + (g->n[1] - 1) * i2);
}

/* Element matrix + load for element (i0,i1[,i2]); idx receives the node numbers.
 * mode 0 = thermal (3D scales k by 1e-6, therm3d.cpp:215; load from heat density),
 * mode 1 = electrical (3D scales by 1e-6, electr3d.cpp:312-313; no load vector). */
static int element_system(const orc_grid* g, int mode, const double* kh, const double* kv, const double* heat,
                          int64_t i0, int64_t i1, int64_t i2, int64_t* idx, double K[8][8], double* F) {
    int64_t e = eidx(g, i0, i1, i2);
    double dx = g->ax[0][i0 + 1] - g->ax[0][i0];
    double dy = g->ax[1][i1 + 1] - g->ax[1][i1];
    if (g->dim == 2) {
        double K2[4][4];
        elem_K2(kh[e], kv[e], dx, dy, K2);
        double r = g->cyl ? 0.5 * (g->ax[0][i0] + g->ax[0][i0 + 1]) : 1.;
        int64_t ll = i0 * g->s[0] + i1 * g->s[1];
        idx[0] = ll;                      /* 1: lo,lo */
        idx[1] = ll + g->s[0];            /* 2: up,lo */
        idx[2] = ll + g->s[0] + g->s[1];  /* 3: up,up */
        idx[3] = ll + g->s[1];            /* 4: lo,up */
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) K[a][b] = r * K2[a][b];
        /* therm2d.cpp:220 (Cartesian), :346 (cylindrical: x r) */
        double f = (mode == 0 && heat) ? 0.25e-12 * r * dx * dy * heat[e] : 0.;
        for (int a = 0; a < 4; ++a) F[a] = f;
        return 4;
    } else {
        double dz = g->ax[2][i2 + 1] - g->ax[2][i2];
        double kxy = kh[e] * 1e-6, kz = kv[e] * 1e-6; /* W/m -> W/um */
        elem_K3(kxy, kxy, kz, dx, dy, dz, K);
        int64_t b = i0 * g->s[0] + i1 * g->s[1] + i2 * g->s[2];
        for (int a = 0; a < 8; ++a)
            idx[a] = b + ((a & 1) ? g->s[0] : 0) + ((a & 2) ? g->s[1] : 0) + ((a & 4) ? g->s[2] : 0);
        /* therm3d.cpp:223 */
        double f = (mode == 0 && heat) ? 0.125e-18 * dx * dy * dz * heat[e] : 0.;
        for (int a = 0; a < 8; ++a) F[a] = f;
        return 8;
    }
}

/* Assemble A (lower triangle, therm3d.cpp:270-275) and B, then apply Dirichlet
 * conditions in list order (matrix.hpp:111-118). Returns 0, or -1 if the
 * storage cannot hold an entry. */
int orc_assemble(const orc_grid* g, int mode, const double* kh, const double* kv, const double* heat,
                 int kind, int64_t ld, int64_t kd, int ndiag, const int64_t* offsets, double* data, double* B,
                 const int64_t* dnodes, const double* dvals, int64_t nd) {
    orc_mat A = {kind, orc_node_count(g), ld, kd, ndiag, offsets, data};
    int64_t size = (kind == ST_BAND) ? A.rank * (ld + 1) : A.rank * ndiag;
    memset(data, 0, sizeof(double) * size);
    memset(B, 0, sizeof(double) * A.rank);
    int64_t m2 = (g->dim == 3) ? g->n[2] - 1 : 1;
    for (int64_t i2 = 0; i2 < m2; ++i2)
        for (int64_t i1 = 0; i1 < g->n[1] - 1; ++i1)
            for (int64_t i0 = 0; i0 < g->n[0] - 1; ++i0) {
                int64_t idx[8];
                double K[8][8], F[8];
                int nn = element_system(g, mode, kh, kv, heat, i0, i1, i2, idx, K, F);
                for (int i = 0; i < nn; ++i) {
                    for (int j = 0; j <= i; ++j) {
                        double* p = mat_ref(&A, idx[i], idx[j]);
                        if (!p) return -1;
                        *p += K[i][j];
                    }
                    B[idx[i]] += F[i];
                }
            }
    /* Dirichlet: BandMatrix::setBC (matrix.hpp:132-145) / SparseBandMatrix::setBC
     * (iterative_matrix.hpp:462-485): symmetric row/column elimination. */
    for (int64_t q = 0; q < nd; ++q) {
        int64_t r = dnodes[q];
        double val = dvals[q];
        B[r] = val;
        *mat_ref(&A, r, r) = 1.;
        if (kind == ST_BAND) {
            int64_t start = (r > kd) ? r - kd : 0;
            int64_t end = (r + kd < A.rank) ? r + kd + 1 : A.rank;
            for (int64_t c = start; c < end; ++c) {
                if (c == r) continue;
                double* p = mat_ref(&A, r, c);
                B[c] -= *p * val;
                *p = 0.;
            }
        } else {
            for (int d = ndiag - 1; d > 0; --d) {
                int64_t c = r - offsets[d];
                if (c >= 0) {
                    double* p = data + c + A.rank * d;
                    B[c] -= *p * val;
                    *p = 0.;
                }
            }
            for (int d = 1; d < ndiag; ++d) {
                int64_t c = r + offsets[d];
                if (c < A.rank) {
                    double* p = data + r + A.rank * d;
                    B[c] -= *p * val;
                    *p = 0.;
                }
            }
        }
    }
    return 0;
}

/* y = A x for the symmetric diagonal storage (iterative_matrix.hpp:420-433). */
void orc_dia_mult(int64_t rank, int ndiag, const int64_t* offsets, const double* data, const double* x, double* y) {
    for (int64_t r = 0; r < rank; ++r) y[r] = data[r] * x[r];
    for (int d = 1; d < ndiag; ++d) {
        const double* dd = data + rank * d;
        int64_t off = offsets[d];
        for (int64_t r = 0; r + off < rank; ++r) {
            int64_t c = r + off;
            y[r] += dd[r] * x[c];
            y[c] += dd[r] * x[r];
        }
    }
}

/* y = A x for the LAPACK 'L' band storage (the dsbmv call of cholesky_matrix.hpp:113-115). */
void orc_band_mult(int64_t rank, int64_t ld, int64_t kd, const double* ab, const double* x, double* y) {
    for (int64_t r = 0; r < rank; ++r) y[r] = 0.;
    for (int64_t c = 0; c < rank; ++c) {
        const double* col = ab + ld * c;
        y[c] += col[c] * x[c];
        int64_t end = (c + kd < rank) ? c + kd : rank - 1;
        for (int64_t r = c + 1; r <= end; ++r) {
            y[r] += col[r] * x[c];
            y[c] += col[r] * x[r];
        }
    }
}

/* Right-looking band Cholesky A = L L^T in place on the 'L' band layout
 * (the factorization dpbtrf computes; Golub & Van Loan, Matrix Computations,
 * band Cholesky). Returns 0, or k+1 if the leading minor of order k+1 is not
 * positive definite (the info convention of cholesky_matrix.hpp:96-100). */
int64_t orc_band_cholesky(int64_t rank, int64_t ld, int64_t kd, double* ab) {
    for (int64_t k = 0; k < rank; ++k) {
        double* ck = ab + ld * k;
        double d = ck[k];
        if (!(d > 0.)) return k + 1;
        d = sqrt(d);
        ck[k] = d;
        int64_t end = (k + kd < rank) ? k + kd : rank - 1;
        double inv = 1. / d;
        for (int64_t r = k + 1; r <= end; ++r) ck[r] *= inv;
        for (int64_t j = k + 1; j <= end; ++j) {
            double ljk = ck[j];
            if (ljk == 0.) continue;
            double* cj = ab + ld * j;
            for (int64_t r = j; r <= end; ++r) cj[r] -= ck[r] * ljk;
        }
    }
    return 0;
}

/* Solve L L^T x = b in place (dpbtrs, cholesky_matrix.hpp:103-111). */
void orc_band_solve(int64_t rank, int64_t ld, int64_t kd, const double* ab, double* b) {
    for (int64_t k = 0; k < rank; ++k) {
        const double* ck = ab + ld * k;
        double v = b[k] / ck[k];
        b[k] = v;
        int64_t end = (k + kd < rank) ? k + kd : rank - 1;
        for (int64_t r = k + 1; r <= end; ++r) b[r] -= ck[r] * v;
    }
    for (int64_t k = rank - 1; k >= 0; --k) {
        const double* ck = ab + ld * k;
        int64_t end = (k + kd < rank) ? k + kd : rank - 1;
        double s = b[k];
        for (int64_t r = k + 1; r <= end; ++r) s -= ck[r] * b[r];
        b[k] = s / ck[k];
    }
}

/* IC(0) on the symmetric diagonal storage: A ~ L D^-1 L^T restricted to the
 * stored pattern, with D the pivots (the IC preconditioner family of NSPCG
 * ic2/pfact2, nspcg.f:1697,6968). Stores pivots in piv[rank] and the strictly
 * lower factor values in lfac (same layout as data). Nonpositive pivots are
 * replaced by the original diagonal (NSPCG reports them as warning ier=5).
 * Returns the number of replaced pivots. */
int64_t orc_ic0_factor(int64_t rank, int ndiag, const int64_t* offsets, const double* data, double* piv, double* lfac) {
    int64_t bad = 0;
    memcpy(lfac, data, sizeof(double) * rank * ndiag);
    for (int64_t i = 0; i < rank; ++i) piv[i] = data[i];
    /* row-oriented: pivot i final once all c<i done; update later diagonal
     * entries and off-diagonals within the pattern (fill dropped). */
    for (int64_t c = 0; c < rank; ++c) {
        double p = piv[c];
        if (!(p > 0.)) { p = data[c] > 0. ? data[c] : 1.; piv[c] = p; ++bad; }
        double ip = 1. / p;
        /* for every pair (r1 = c + o1, r2 = c + o2), o1 <= o2 both in pattern:
         * a(r2, r1) -= l(r1,c) l(r2,c) / p if (r2 - r1) is a stored offset */
        for (int d1 = 1; d1 < ndiag; ++d1) {
            int64_t r1 = c + offsets[d1];
            if (r1 >= rank) continue;
            double l1 = lfac[c + rank * d1];
            if (l1 == 0.) continue;
            piv[r1] -= l1 * l1 * ip;
            for (int d2 = d1 + 1; d2 < ndiag; ++d2) {
                int64_t r2 = c + offsets[d2];
                if (r2 >= rank) continue;
                double l2 = lfac[c + rank * d2];
                if (l2 == 0.) continue;
                int64_t off = offsets[d2] - offsets[d1];
                for (int d3 = 1; d3 < ndiag; ++d3)
                    if (offsets[d3] == off) { lfac[r1 + rank * d3] -= l1 * l2 * ip; break; }
            }
        }
    }
    return bad;
}

/* z = (L D^-1 L^T)^-1 r with unit scaling folded: solve (D + Ls) D^-1 (D + Ls^T) z = r */
static void ic0_apply(int64_t rank, int ndiag, const int64_t* offsets, const double* piv, const double* lfac,
                      const double* r, double* z, double* w) {
    /* forward: (D + Ls) w = r, Ls(r2,c) = lfac at (c, d) with r2 = c + off */
    for (int64_t i = 0; i < rank; ++i) w[i] = r[i];
    for (int64_t c = 0; c < rank; ++c) {
        w[c] /= piv[c];
        double v = w[c];
        for (int d = 1; d < ndiag; ++d) {
            int64_t r2 = c + offsets[d];
            if (r2 < rank) w[r2] -= lfac[c + rank * d] * v;
        }
    }
    /* w now holds D^-1 (D+Ls)^-1 r scaled: w = (D+Ls)^-1 r / ... ; back: (I + D^-1 Ls^T) z = w */
    for (int64_t c = rank - 1; c >= 0; --c) {
        double s = w[c];
        for (int d = 1; d < ndiag; ++d) {
            int64_t r2 = c + offsets[d];
            if (r2 < rank) s -= lfac[c + rank * d] * z[r2] / piv[c];
        }
        z[c] = s;
    }
}

static double dot(int64_t n, const double* a, const double* b) {
    double s = 0.;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Preconditioned CG in the order of NSPCG itcg (nspcg.f:9217-9330):
 *   r = b - A u; loop { z = Q^-1 r; rz = (r,z) (<=0 -> -7); stop test;
 *   beta = rz/rz_old; p = z + beta p; Ap; alpha = rz/(p,Ap) (<=0 -> -6);
 *   u += alpha p; r -= alpha Ap }
 * The stop test is the true relative residual ||b - A u|| / bnorm_ref <= rtol
 * (bnorm_ref <= 0: ||b||), recursive r confirmed by an explicit residual.
 * precond: 0 = none, 1 = Jacobi (pjac), 2 = IC(0).
 * Returns iterations >= 0, or -6/-7 on breakdown, or -1 when maxit is reached
 * (iterations still written to *iters). */
int64_t orc_dia_pcg(int64_t rank, int ndiag, const int64_t* offsets, const double* data, const double* b, double* u,
                    int precond, double rtol, int64_t maxit, double bnorm_ref, double* rel_out, int64_t* iters) {
    double* r = malloc(sizeof(double) * rank);
    double* z = malloc(sizeof(double) * rank);
    double* p = malloc(sizeof(double) * rank);
    double* q = malloc(sizeof(double) * rank);
    double *piv = NULL, *lfac = NULL, *w = NULL;
    if (precond == 2) {
        piv = malloc(sizeof(double) * rank);
        lfac = malloc(sizeof(double) * rank * ndiag);
        w = malloc(sizeof(double) * rank);
        orc_ic0_factor(rank, ndiag, offsets, data, piv, lfac);
    }
    double bn = bnorm_ref > 0. ? bnorm_ref : sqrt(dot(rank, b, b));
    if (bn == 0.) bn = 1.;
    orc_dia_mult(rank, ndiag, offsets, data, u, r);
    for (int64_t i = 0; i < rank; ++i) r[i] = b[i] - r[i];
    double rz = 0., rz_old = 0.;
    int64_t it = 0, status = 0;
    for (;;) {
        if (precond == 1)
            for (int64_t i = 0; i < rank; ++i) z[i] = r[i] / data[i];
        else if (precond == 2)
            ic0_apply(rank, ndiag, offsets, piv, lfac, r, z, w);
        else
            memcpy(z, r, sizeof(double) * rank);
        rz_old = rz;
        rz = dot(rank, r, z);
        double rn = sqrt(dot(rank, r, r));
        if (rn / bn <= rtol) {
            /* confirm on the true residual */
            orc_dia_mult(rank, ndiag, offsets, data, u, q);
            double t = 0.;
            for (int64_t i = 0; i < rank; ++i) { double d = b[i] - q[i]; t += d * d; }
            if (sqrt(t) / bn <= rtol) { *rel_out = sqrt(t) / bn; status = it; break; }
            for (int64_t i = 0; i < rank; ++i) r[i] = b[i] - q[i];
            continue;
        }
        if (rz <= 0.) { status = -7; break; }
        if (it >= maxit) { *rel_out = rn / bn; status = -1; break; }
        /* first direction p = z: p is uninitialised memory then, and 0 * (NaN garbage) is NaN */
        double beta = (it == 0) ? 0. : rz / rz_old;
        if (it == 0)
            memcpy(p, z, sizeof(double) * rank);
        else
            for (int64_t i = 0; i < rank; ++i) p[i] = z[i] + beta * p[i];
        orc_dia_mult(rank, ndiag, offsets, data, p, q);
        double pap = dot(rank, p, q);
        if (pap <= 0.) { status = -6; break; }
        double alpha = rz / pap;
        for (int64_t i = 0; i < rank; ++i) { u[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        ++it;
    }
    *iters = it;
    free(r); free(z); free(p); free(q); free(piv); free(lfac); free(w);
    return status;
}

/* Matrix-free apply of the eliminated operator, element by element:
 * y = M A0 (M x) + (I - M) x with M the free-node mask (dmask[n] = 1 for Dirichlet).
 * This is the operator the assembled system of orc_assemble represents. */
void orc_apply(const orc_grid* g, int mode, const double* kh, const double* kv, const uint8_t* dmask, const double* x,
               double* y) {
    int64_t N = orc_node_count(g);
    for (int64_t i = 0; i < N; ++i) y[i] = 0.;
    int64_t m2 = (g->dim == 3) ? g->n[2] - 1 : 1;
    for (int64_t i2 = 0; i2 < m2; ++i2)
        for (int64_t i1 = 0; i1 < g->n[1] - 1; ++i1)
            for (int64_t i0 = 0; i0 < g->n[0] - 1; ++i0) {
                int64_t idx[8];
                double K[8][8], F[8];
                int nn = element_system(g, mode, kh, kv, NULL, i0, i1, i2, idx, K, F);
                for (int a = 0; a < nn; ++a) {
                    if (dmask[idx[a]]) continue;
                    double s = 0.;
                    for (int b = 0; b < nn; ++b)
                        if (!dmask[idx[b]]) s += K[a][b] * x[idx[b]];
                    y[idx[a]] += s;
                }
            }
    for (int64_t i = 0; i < N; ++i)
        if (dmask[i]) y[i] = x[i];
}

/* Junction conductivity update of the Shockley solver (electr2d.cpp:243-264,
 * electr3d.cpp:246-273) with the law of BetaSolver::activeCond (beta.hpp:43-46):
 *   U = mean over the element's corner columns of V(top) - V(bottom)
 *   jy = 0.1 sigma_yy U / d;  sigma = 10 |jy| d beta / ln(1e7 |jy| / js + 1)
 *   convergence 1 = STABLE: average with the old value; clamp at 1e-16.
 * elems: element indices of junction elements; cols: for each, the node
 * numbers of its bottom and top corner columns (2 per side in 2D, 4 in 3D),
 * laid out as [ncorner bottom, ncorner top]. */
void orc_junction_update(int64_t nje, int ncorner, const int64_t* elems, const int64_t* corners, const double* V,
                         double height, double beta, double js, int stable, double* sig_v /* per element, in/out */) {
    for (int64_t q = 0; q < nje; ++q) {
        const int64_t* c = corners + q * 2 * ncorner;
        double U = 0.;
        for (int k = 0; k < ncorner; ++k) U += V[c[ncorner + k]] - V[c[k]];
        U /= ncorner;
        int64_t e = elems[q];
        double jy = 0.1 * sig_v[e] * U / height;
        jy = fabs(jy);
        double cond = 10. * jy * height * beta / log(1e7 * jy / js + 1.);
        if (stable) cond = 0.5 * (sig_v[e] + cond);
        if (isnan(cond) || fabs(cond) < 1e-16) cond = 1e-16;
        sig_v[e] = cond;
    }
}

/* Element current densities j = -sigma grad V in kA/cm^2 (electr2d.cpp:405-409,
 * electr3d.cpp:399-411) and Joule heat sigma |grad V|^2 in W/m^3
 * (electr2d.cpp:457-461, electr3d.cpp:459-475). Output per element:
 * cur[dim*e + a], heat[e] (heat may be NULL). */
void orc_currents(const orc_grid* g, const double* sh, const double* sv, const double* V, double* cur, double* heat) {
    int64_t m2 = (g->dim == 3) ? g->n[2] - 1 : 1;
    for (int64_t i2 = 0; i2 < m2; ++i2)
        for (int64_t i1 = 0; i1 < g->n[1] - 1; ++i1)
            for (int64_t i0 = 0; i0 < g->n[0] - 1; ++i0) {
                int64_t e = eidx(g, i0, i1, i2);
                double dx = g->ax[0][i0 + 1] - g->ax[0][i0];
                double dy = g->ax[1][i1 + 1] - g->ax[1][i1];
                int64_t b = i0 * g->s[0] + i1 * g->s[1] + (g->dim == 3 ? i2 * g->s[2] : 0);
                if (g->dim == 2) {
                    double ll = V[b], lu = V[b + g->s[0]], ul = V[b + g->s[1]], uu = V[b + g->s[0] + g->s[1]];
                    double gx = (-ll + lu - ul + uu) / dx, gy = (-ll - lu + ul + uu) / dy;
                    cur[2 * e] = sh[e] * (-0.05 * gx);
                    cur[2 * e + 1] = sv[e] * (-0.05 * gy);
                    if (heat) {
                        double hx = 0.5e6 * gx, hy = 0.5e6 * gy;
                        heat[e] = sh[e] * hx * hx + sv[e] * hy * hy;
                    }
                } else {
                    double dz = g->ax[2][i2 + 1] - g->ax[2][i2];
                    double v[8];
                    for (int a = 0; a < 8; ++a)
                        v[a] = V[b + ((a & 1) ? g->s[0] : 0) + ((a & 2) ? g->s[1] : 0) + ((a & 4) ? g->s[2] : 0)];
                    double sx = 0., sy = 0., sz = 0.;
                    for (int a = 0; a < 8; ++a) {
                        sx += (a & 1) ? v[a] : -v[a];
                        sy += (a & 2) ? v[a] : -v[a];
                        sz += (a & 4) ? v[a] : -v[a];
                    }
                    cur[3 * e] = -0.025 * sh[e] * sx / dx;
                    cur[3 * e + 1] = -0.025 * sh[e] * sy / dy;
                    cur[3 * e + 2] = -0.025 * sv[e] * sz / dz;
                    if (heat) {
                        double hx = -0.25e6 * sx / dx, hy = -0.25e6 * sy / dy, hz = -0.25e6 * sz / dz;
                        heat[e] = sh[e] * hx * hx + sh[e] * hy * hy + sv[e] * hz * hz;
                    }
                }
            }
}

double orc_wtime(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
