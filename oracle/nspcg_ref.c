/* Driver of the reference's own NSPCG (TEST INFRASTRUCTURE ONLY: built by oracle/Makefile
 * from the reference sources into oracle/_ref/, called by tests and by the CPU baseline
 * leg of bench.py; never by the product).
 *
 * It calls NSPCG the way SparseMatrix::solverhs does (plask/common/fem/iterative_matrix.hpp
 * :141-268) for the symmetric diagonal storage of SparseBandMatrix (:347-486): defaults
 * from dfault, then nstore = 2, itmax = maxit, ipropa = 0, ifact = 1 on a new matrix,
 * zeta = maxerr, ns1/ns2/lvfill/ltrunc/ndeg/omega at PLaSK's defaults (:80-86), level = -1
 * (no output), kblsz / nbl2d from the grid; the workspace sizes of :175-190, grown when
 * NSPCG asks for more (ier -2 / -3, :252-264).  Preconditioner ic2 (PLaSK's default,
 * PRECOND_IC) or jac2 (PRECOND_JAC), accelerator cg (ACCEL_CG).
 *
 * The reference builds NSPCG from nspcg.f with gfortran, which this image lacks; its f2c
 * translation nspcg.c (same algorithm, Fortran integers as C long) is built instead. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long integer;  /* f2c's integer (extlib/libf2c/f2c.h:10) */
typedef double doublereal;
typedef int (*S_fp)();
typedef int (*U_fp)();

extern int nspcg_(S_fp precon, U_fp accel, integer* ndimm, integer* mdimm, integer* nn, integer* maxnzz,
                  doublereal* coef, integer* jcoef, integer* p, integer* ip, doublereal* u, doublereal* ubar,
                  doublereal* rhs, doublereal* wksp, integer* iwksp, integer* nw, integer* inw, integer* iparm,
                  doublereal* rparm, integer* ier);
extern int dfault_(integer* iparm, doublereal* rparm);
extern int ic2_();
extern int jac2_();
extern int cg_();

/* Solve with NSPCG.  coef: rank x ndiag symmetric diagonal storage (column d holds
 * A(c, c + offsets[d]) at c), copied so NSPCG may work on it; u: start vector in,
 * solution out.  precond 0 = jac2, 1 = ic2.  minor / major: node strides of the grid
 * (kblsz = minor - 1, nbl2d = major - minor - 1; 2D: major = 0 gives nbl2d = minor - 1).
 * Outputs: iterations (iparm itmax on exit), the final stopping-test value (rparm zeta),
 * NSPCG's own factorisation and iteration times.  Returns ier. */
long orc_ref_nspcg(long rank, int ndiag, const int64_t* offsets, const double* coef_in, const double* rhs_in,
                   double* u, int precond, long maxit, double zeta, long minor, long major, long* iters,
                   double* zeta_out, double* t_fact, double* t_iter) {
    integer iparm[30], ier = 0;
    doublereal rparm[30];
    memset(iparm, 0, sizeof(iparm));
    memset(rparm, 0, sizeof(rparm));
    dfault_(iparm, rparm);
    iparm[11] = 2;          /* nstore: symmetric diagonal storage */
    iparm[1] = maxit;       /* itmax */
    iparm[17] = 0;          /* ipropa */
    iparm[14] = 1;          /* ifact: factorise this matrix */
    rparm[0] = zeta;        /* zeta */
    iparm[8] = 5;           /* ns1 */
    iparm[9] = 100000;      /* ns2 */
    iparm[15] = 0;          /* lvfill */
    iparm[16] = 0;          /* ltrunc */
    iparm[24] = 1;          /* ndeg */
    rparm[8] = 1.0;         /* omega */
    iparm[10] = 0;          /* ns3 */
    iparm[18] = minor - 1;  /* kblsz */
    iparm[19] = major ? major - minor - 1 : minor - 1; /* nbl2d */
    iparm[2] = -1;          /* level */
    integer n = rank, ndim = rank, mdim = ndiag, maxnz = ndiag;
    integer* jcoef = malloc(sizeof(integer) * ndiag);
    for (int d = 0; d < ndiag; ++d) jcoef[d] = (integer)offsets[d];
    double* coef = malloc(sizeof(double) * (size_t)rank * ndiag);
    memcpy(coef, coef_in, sizeof(double) * (size_t)rank * ndiag);
    double* rhs = malloc(sizeof(double) * (size_t)rank);
    memcpy(rhs, rhs_in, sizeof(double) * (size_t)rank);
    long kb = iparm[18] > 1 ? iparm[18] : 1;
    integer nw = 3 * rank + 2 * maxit + rank * maxnz + kb;
    integer inw = maxnz + (2 * rank > maxnz * maxnz + maxnz ? 2 * rank : maxnz * maxnz + maxnz);
    double* wksp = malloc(sizeof(double) * (size_t)nw);
    integer* iwksp = malloc(sizeof(integer) * (size_t)inw);
    for (;;) {
        nspcg_(precond ? (S_fp)ic2_ : (S_fp)jac2_, (U_fp)cg_, &ndim, &mdim, &n, &maxnz, coef, jcoef, NULL, NULL, u,
               NULL, rhs, wksp, iwksp, &nw, &inw, iparm, rparm, &ier);
        if (ier == -2 && nw) {
            free(wksp);
            wksp = malloc(sizeof(double) * (size_t)nw);
            iparm[14] = 1;
        } else if (ier == -3 && inw) {
            free(iwksp);
            iwksp = malloc(sizeof(integer) * (size_t)inw);
            iparm[14] = 1;
        } else
            break;
    }
    *iters = iparm[1];
    *zeta_out = rparm[0];
    *t_fact = rparm[12]; /* timfac */
    *t_iter = rparm[5];  /* timit */
    free(jcoef);
    free(coef);
    free(rhs);
    free(wksp);
    free(iwksp);
    return ier;
}
