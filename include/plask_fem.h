/*
 * plask_fem.h -- C ABI of the MI355X (gfx950) FEM thermal / electrical solve path.
 *
 * Drop-in boundary for the steady-state FEM path of PLaSK's `thermal.static` and
 * `electrical.shockley` solvers.  In the reference the seam is the FemMatrix
 * family (plask/common/fem/matrix.hpp:21-123) driven by each solver's
 * setMatrix()+solve() loop; a matrix-free GPU operator cannot sit behind
 * FemMatrix::operator()(r,c) += v, so this ABI replaces the pair
 * "setMatrix + FemMatrix::solve" of one compute() loop:
 *
 *   pf_grid_set + pf_coeffs_set + pf_source_set + pf_dirichlet_set
 *       replace ThermalFem3DSolver::setMatrix (solvers/thermal/static/therm3d.cpp:170-279),
 *       ThermalFem2DSolver<..>::setMatrix (therm2d.cpp:176-298 Cartesian, :301-440 cylindrical),
 *       ElectricalFem2DSolver<..>::setMatrix (solvers/electrical/shockley/electr2d.cpp:234-329),
 *       ElectricalFem3DSolver::setMatrix (electr3d.cpp:239-354)
 *       and FemMatrix::applyBC (plask/common/fem/matrix.hpp:111-118)
 *   pf_solve replaces FemMatrix::solve -> DpbMatrix::factorize/solverhs
 *       (plask/common/fem/cholesky_matrix.hpp:90-111) and
 *       SparseMatrix::solverhs (plask/common/fem/iterative_matrix.hpp:141-339)
 *   pf_apply replaces FemMatrix::mult (matrix.hpp:89; cholesky_matrix.hpp:113-115;
 *       iterative_matrix.hpp:341-344)
 *
 * Conventions
 *   - Status codes, never exceptions: 0 = OK, < 0 = error (pf_last_error has the text).
 *   - Node numbering is natural order, axis 0 fastest: node = i0 + n0*(i1 + n1*i2).
 *     Element numbering likewise over (n0-1, n1-1, n2-1).  (The reference numbers
 *     nodes by the mesh iteration order, plask/mesh/rectilinear3d.cpp:20-23; the
 *     host solver classes translate.)
 *   - Units as in the reference: coordinates in um, conductivities in W/(m K)
 *     (thermal) or S/m (electrical), heat density in W/m^3.  The reference's unit
 *     scalings (x1e-6 in 3D, therm3d.cpp:215 / electr3d.cpp:312-313; x r_mid in
 *     cylindrical 2D, therm2d.cpp:333) are applied inside.
 *   - Host buffers are owned by the caller and copied; device buffers are owned by
 *     the context.  A context is not thread-safe; distinct contexts are independent.
 *     All entry points return after their work is complete on the device.
 */
#ifndef PLASK_FEM_H
#define PLASK_FEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_ctx pf_ctx;

/* status codes */
#define PF_OK 0
#define PF_EINVAL -1   /* bad argument */
#define PF_ESTATE -2   /* call order (e.g. solve before grid) */
#define PF_EDEVICE -3  /* HIP runtime error */
#define PF_ESHAPE -4   /* grid shape outside what the kernels support */
#define PF_ENOCONV -5  /* did not reach the tolerance within maxit (result still written) */
#define PF_EBREAK -6   /* CG breakdown: (p,Ap) <= 0 or (r,z) <= 0 -- matrix not positive definite */
#define PF_ENONFINITE -8

/* geometry */
#define PF_GEOM_CARTESIAN 0
#define PF_GEOM_CYLINDRICAL 1 /* 2D only: axis 0 is the radius */

/* equation */
#define PF_THERMAL 0
#define PF_ELECTRICAL 1

/* precision of a solve */
#define PF_FP64 0         /* FP64 vectors and operator */
#define PF_FP32_REFINE 1  /* FP32 operator + inner PCG, FP64 outer residual / update */
#define PF_FP32_RELIABLE 2 /* as PF_FP32_REFINE, but one FP32 Krylov run: at each FP64 residual
                              replacement the search direction is kept (reliable updates) */
#define PF_FP64_PREC32 3   /* FP64 vectors, operator and CG; the preconditioner (line solve or
                              multigrid V-cycle) applied in FP32 */

/* start vector */
#define PF_X0_GIVEN 0 /* x_inout holds the start vector (its Dirichlet entries are overwritten) */
#define PF_X0_ZERO 1
#define PF_X0_DEVICE 2 /* start from the solution kept on the device from the previous pf_solve */
#define PF_X0_UNIFORM 3 /* start from init_value everywhere (the reference's inittemp field), on the device */
#define PF_X0_DEVICE_CHECKED 4 /* the device vector (e.g. a learned guess) unless its residual exceeds that of the
                                  uniform init_value start, which is then used instead (norm must be PF_NORM_INIT) */

/* denominator of the relative residual */
#define PF_NORM_INIT 0   /* ||B - A X_init||, X_init = init_value at free nodes + the Dirichlet values:
                            the residual of the reference's start field (inittemp, therm3d.cpp:79;
                            zero potential, electr2d.cpp:176), i.e. ||b|| of the correction system */
#define PF_NORM_B 1      /* ||B|| over the free rows */
#define PF_NORM_B_FULL 2 /* ||B|| including the Dirichlet rows (B exactly as assembled by the reference) */
#define PF_NORM_START 3  /* max(||B - A X_start||, f ||B - A X_init||), f = max(1e-4, 1e-13 / rtol), X_start
                            the start vector of this solve: the start residual reduced by rtol, the way a
                            Picard loop that warm-starts each linear solve from the previous loop's solution
                            matches a direct solve per loop (the floor keeps a converged loop from asking for
                            more than FP64 CG attains, ~3e-14 ||B - A X_init|| on the 3D junction systems:
                            the target never goes below 1e-13 ||B - A X_init||) */

typedef struct {
    int precision;    /* PF_FP64 | PF_FP32_REFINE | PF_FP32_RELIABLE | PF_FP64_PREC32 */
    int precond;      /* 1 = Jacobi (NSPCG 'jac'); 2 = line block-Jacobi along the last (layer) axis
                         (the tridiagonal part of A per column, cf. NSPCG 'ljac'); 3 = multigrid
                         V-cycle (in-plane semicoarsening, damped line smoother; no reference
                         counterpart, see pf_set_multigrid); 0 = default (2) */
    int x0;           /* PF_X0_* */
    int maxit;        /* max CG iterations (sum of inner iterations for PF_FP32_REFINE) */
    double rtol;      /* stop when ||B - A X||_2 <= rtol * (denominator chosen by `norm`) */
    double inner_rtol;/* FP32 forms: relative reduction asked of each FP32 run (0: 1e-4 for
                         PF_FP32_REFINE, 1e-2 for PF_FP32_RELIABLE) */
    int max_outer;    /* PF_FP32_REFINE: max refinement steps */
    int check_every;  /* iterations between host convergence checks (0 = default 32) */
    int norm;         /* PF_NORM_* */
    double init_value;/* X_init for PF_NORM_INIT */
} pf_solve_opts;

typedef struct {
    int iters;          /* CG iterations (inner iterations summed for PF_FP32_REFINE) */
    int outer;          /* refinement steps */
    double rel_resid;   /* true ||B - A X|| / denominator at exit, computed in FP64 */
    double bnorm;       /* the denominator */
    double seconds;     /* wall time of the solve on the device */
    int status;         /* PF_OK or an error code */
    int start;          /* PF_X0_DEVICE_CHECKED: 0 = the device vector was used, 1 = the uniform start */
    int fallback;       /* FP32 forms: 1 when a refinement step failed to halve the FP64 residual (the FP32
                           correction no longer contracts: kappa(A) eps_fp32 ~ 1) and the solve finished with
                           FP64 CG; `outer` counts the FP32 steps and the FP64 runs */
} pf_solve_stats;

/* context ------------------------------------------------------------------ */
int pf_ctx_create(pf_ctx** out, int device, int flags);
void pf_ctx_destroy(pf_ctx* ctx);
const char* pf_last_error(const pf_ctx* ctx);

/* Rectilinear grid.  dim 2: (ax0,n0) x (ax1,n1); dim 3 adds (ax2,n2).  Replaces the
 * mesh part of FemSolverWithMaskedMesh::getMatrix (plask/common/fem/fem_solver.hpp:196-241)
 * for the full (non-masked) mesh. */
int pf_grid_set(pf_ctx* ctx, int dim, int geom, const double* ax0, int64_t n0, const double* ax1, int64_t n1,
                const double* ax2, int64_t n2);

/* Per-element conductivities: in-plane (c00 of Tensor2) and vertical (c11).  3D uses
 * kx = ky = in-plane (therm3d.cpp:213-216, electr3d.cpp:310-314).  equation = PF_THERMAL
 * or PF_ELECTRICAL selects nothing numerically (both use the same element matrix) but
 * is recorded for the load vector (electrical has none, electr2d.cpp:268-317). */
int pf_coeffs_set(pf_ctx* ctx, int equation, const double* k_inplane, const double* k_vert, int64_t nelem);

/* Per-element heat density (W/m^3) for the load vector (therm2d.cpp:220, :346;
 * therm3d.cpp:223).  NULL = zero. */
int pf_source_set(pf_ctx* ctx, const double* q_elem, int64_t nelem);

/* Dirichlet nodes and values (BandMatrix::setBC, matrix.hpp:132-145; first kind
 * boundary conditions).  A node listed twice takes the last value. */
int pf_dirichlet_set(pf_ctx* ctx, const int64_t* nodes, const double* vals, int64_t n);

/* Nodal load added to the element load vector before the Dirichlet elimination: the
 * boundary heat-flux and radiation terms of setMatrix (therm2d.cpp:139-173 with
 * :232-240, :256-265; cylindrical :358-368, :394-406; therm3d.cpp:141-168 with :242-249,
 * :261-268), evaluated by the caller.  Repeated nodes are summed; n = 0 clears it.  Kept until replaced. */
int pf_load_set(pf_ctx* ctx, const int64_t* nodes, const double* vals, int64_t n);

/* Boundary terms of the stiffness matrix: the convection entries of setMatrix
 * (setBoundaries with the convection K functions, therm2d.cpp:242-254, cylindrical
 * :370-392, therm3d.cpp:250-259), evaluated by the caller as the reference adds them,
 * A(row, col) += val for col <= row (an off-diagonal entry stands for both triangles).
 * Repeated entries are summed; nnz = 0 clears them.  They are part of the operator of
 * pf_solve, pf_apply and every residual, and take part in the Dirichlet elimination like
 * the element entries; the preconditioners see their diagonal.  Not combined with a
 * transient operator (pf_capacity_set). */
int pf_boundary_matrix_set(pf_ctx* ctx, const int64_t* rows, const int64_t* cols, const double* vals, int64_t nnz);

/* Linear interpolation of the current solution (node values) at npts points
 * (pts: npts x dim, row-major; um), the providers' lazy interpolation from the solver
 * mesh to a destination mesh (getTemperatures therm2d.cpp:543-552, therm3d.cpp:387-395;
 * getVoltage electr2d.cpp:507-517); points outside the grid get `fill`. */
int pf_interp_nodes(pf_ctx* ctx, const double* pts, int64_t npts, double fill, double* out);

/* Solve the eliminated system A X = B.  x_inout: host array of N node values
 * (start vector in, solution out); may be NULL with PF_X0_ZERO / PF_X0_DEVICE to
 * leave the solution on the device (pf_get_solution copies it out). */
int pf_solve(pf_ctx* ctx, const pf_solve_opts* opts, double* x_inout, pf_solve_stats* stats);

/* Copy the device solution of the last pf_solve to the host (N values). */
int pf_get_solution(pf_ctx* ctx, double* x_out);

/* y = A x on device buffers of N values (FP64 when precision = PF_FP64, FP32 when
 * PF_FP32_REFINE), A the Dirichlet-eliminated operator. */
int pf_apply(pf_ctx* ctx, int precision, const void* x_dev, void* y_dev);

/* Host-buffer convenience: y = A x (FP64). */
int pf_apply_host(pf_ctx* ctx, const double* x, double* y);

/* Host-buffer test hook: z = M^-1 r for one preconditioner (1 Jacobi, 2 line, 3 one
 * multigrid V-cycle) in the given precision (r rounded to FP32 for the FP32 forms).
 * Overwrites the device residual and direction vectors of the last solve. */
int pf_precond_apply_host(pf_ctx* ctx, int precision, int precond, const double* r_in, double* z_out);

/* The eliminated load vector B (host, N values). */
int pf_rhs_host(pf_ctx* ctx, double* b_out);

/* Device pointer of the solution vector (FP64, N values) kept by the context. */
void* pf_solution_device(pf_ctx* ctx);

/* Number of nodes / elements of the current grid. */
int64_t pf_num_nodes(const pf_ctx* ctx);
int64_t pf_num_elements(const pf_ctx* ctx);

/* Kernel classes of pf_kernel_times */
#define PF_KCAT_APPLY 0       /* fused operator apply + (p, Ap) */
#define PF_KCAT_PRECOND 1     /* residual update + preconditioner + (r, z): the line / Jacobi kernel or a whole V-cycle */
#define PF_KCAT_MG_RESIDUAL 2 /* V-cycle on the solve grid: smoother residual applies */
#define PF_KCAT_MG_SMOOTH 3   /* V-cycle on the solve grid: damped line sweeps */
#define PF_KCAT_MG_TRANSFER 4 /* V-cycle: restriction / prolongation between the solve grid and level 1 */
#define PF_KCAT_MG_COARSE 5   /* V-cycle: all work on the coarse levels (one record per V-cycle) */
#define PF_KCAT_COUNT 6

/* Kernels of one PCG iteration, for pf_kernel_stats (each id is one kernel function of
 * this library; the multigrid ones run on every level):
 *   PF_K_APPLY     fused operator apply (k_apply3d / k_apply3r / k_apply2d, MODE_OP, p and dot fused)
 *   PF_K_RESIDUAL  smoother residual r = b - A x (k_apply3d / k_apply2d, MODE_RESG)
 *   PF_K_SMOOTH    damped line sweep (k_mg_smooth)
 *   PF_K_RESTRICT  restriction (k_mg_restrict), level = the finer level
 *   PF_K_PROLONG   prolongation (k_mg_prolong_add), level = the finer level
 *   PF_K_LINE      line preconditioner update (k_cg_part, or k_cg_line when sequential)
 *   PF_K_JACOBI    Jacobi update (k_cg_update)
 *   PF_K_TO_F32, PF_K_FROM_F32  precision conversions around an FP32 preconditioner */
#define PF_K_APPLY 0
#define PF_K_RESIDUAL 1
#define PF_K_SMOOTH 2
#define PF_K_RESTRICT 3
#define PF_K_PROLONG 4
#define PF_K_LINE 5
#define PF_K_JACOBI 6
#define PF_K_TO_F32 7
#define PF_K_FROM_F32 8

typedef struct {
    int kernel;       /* PF_K_* */
    int level;        /* multigrid level (0 = the solve grid) */
    int value_bytes;  /* 8 (FP64) or 4 (FP32) */
    int launches;     /* timed launches */
    double seconds;   /* summed duration of those launches (HIP events around each on the context stream) */
    double bytes;     /* summed algorithmic HBM bytes: every value read or written once, two conductivities
                         per element, the Dirichlet mask as one bit per node */
} pf_kernel_stat;

/* Per-kernel statistics of the iterations of the last pf_solve run with profiling on
 * (pf_set_profile): one entry per (kernel, level, precision) that ran.  Copies up to nmax
 * entries; returns the number of entries (or < 0 on error).  For benchmarks. */
int pf_kernel_stats(const pf_ctx* ctx, pf_kernel_stat* out, int nmax);

/* Average duration (s) of each kernel class (PF_KCAT_*, n <= 8 entries) over the
 * iterations of the last pf_solve run with profiling on (pf_set_profile), measured with
 * HIP events around each launch on the context stream; counts[] receives the number of
 * timed launches per class.  For benchmarks. */
int pf_kernel_times(pf_ctx* ctx, double* seconds, int64_t* counts, int n);

/* Launch shape of the operator kernels (for tuning): 3D rows = thread rows per block
 * (rows * n0 <= 1024, one of them a halo row), planes = planes per block along axis 2;
 * 2D planes = rows of axis 1 per block. */
int pf_set_launch(pf_ctx* ctx, int rows, int planes);

/* Line preconditioner schedule (the preconditioner itself, the exact tridiagonal
 * solve along whole lines of the layer axis, does not change): seg = 4, 8, 16 or
 * 32 solves each line as segments of seg planes, one thread each, coupled through
 * the Schur complement of the segment separators (FP64 uses at most 16), on every
 * multigrid level; seg = 0 restores the default (8 FP64, 16 FP32 on the solve grid;
 * the multigrid coarse levels take 8, and 4 on levels of at most 512 columns, where
 * the separators stay at most 64); seg < 0 selects the sequential one-thread-per-line
 * kernel.  No counterpart in the reference (a tuning hook). */
int pf_set_line_segment(pf_ctx* ctx, int seg);

/* Multigrid preconditioner (precond = 3): levels coarsen the in-plane axes (3D: axes 0
 * and 1; 2D: axis 0) by two (an axis stops at 3 nodes), keep every plane of the layer
 * axis, and use the same
 * Q1 operator with area-weighted mean conductivities; the smoother is the exact line
 * solve along the layer axis damped by omega (0 < omega <= 0.66, default 0.65: the line
 * preconditioned operator has eigenvalues <= 3, so the smoother converges), one sweep
 * before and after each coarse correction and coarse_sweeps (1-64, default 1) on the
 * coarsest level; max_levels limits the level count (0 = no limit).  Needs the
 * partitioned line solve (pf_set_line_segment >= 0).  The levels are built on the first
 * multigrid solve after a grid change.  No counterpart in the reference (NSPCG offers
 * no multigrid); a solver option. */
int pf_set_multigrid(pf_ctx* ctx, int max_levels, int coarse_sweeps, double omega);

/* 3D multigrid: run the grid transfers inside the neighbouring kernels (restriction in a
 * coarse level's first sweep, prolongation in the residual and sweep that follow it; on = 1)
 * or as kernels of their own (0, the default: measured faster).  Same preconditioner
 * either way.  No reference counterpart (tuning hook). */
int pf_set_multigrid_fusion(pf_ctx* ctx, int on);

/* 3D multigrid, levels below the solve grid: residual tiles ordered by in-plane row and
 * line-sweep column groups in contiguous ranges per XCD (hardware block b runs on XCD
 * b % 8), so that consecutive kernels of a level read what the previous one wrote from
 * the same L2 (mode 1, the default); the same on the solve grid's residuals and sweeps
 * too (2, measured slower: the solve grid's vectors do not fit the L2s); or the plain
 * block order (0).  Same arithmetic in every mode.  No reference counterpart (tuning hook). */
int pf_set_multigrid_xcd(pf_ctx* ctx, int mode);

/* Number of multigrid levels built (the solve grid included; 0 before the first
 * multigrid solve). */
int pf_multigrid_levels(const pf_ctx* ctx);

/* 3D operator kernel for one precision (PF_FP64, PF_FP32_REFINE, or -1 for both):
 * kind 0 = tile blocks (shape from pf_set_launch; the default), kind 1 = row blocks,
 * one block per node row and chunk of `planes` >= 4 planes along axis 2 (the default
 * when axis 0 is too long for a tile block).  A tuning hook, no reference counterpart. */
int pf_set_apply(pf_ctx* ctx, int precision, int kind, int planes);

/* Tile-block apply: xcd = 1 (default) gives each of the 8 XCDs a contiguous range of
 * tiles (hardware block b runs on XCD b mod 8), so neighbouring tiles share halo rows in
 * that XCD's L2; 0 = plain block order.  A tuning hook, no reference counterpart. */
int pf_set_block_order(pf_ctx* ctx, int xcd);

/* on = 1: CG iterations are replayed in chunks from hipGraphs; 0: plain launches.
 * Default on, except in a process run under a kernel-tracing profiler (rocprofv3: the
 * environment holds ROCP_TOOL_LIBRARIES or ROCPROF_OUTPUT_PATH), where graph replay
 * is off and a one-line notice goes to stderr.  pf_graphs_enabled reports the setting.
 * No reference counterpart. */
int pf_set_graphs(pf_ctx* ctx, int on);
int pf_graphs_enabled(const pf_ctx* ctx);

/* Planes of node and element data the 3D operator kernel keeps in flight ahead
 * of the plane it computes (1 or 2).  A tuning hook with no reference counterpart. */
int pf_set_prefetch(pf_ctx* ctx, int depth);

/* Device-to-device streaming copy of `bytes` bytes (read + write), timed over
 * `reps` launches: the achievable HBM bandwidth and the calibration kernel of
 * hardware traffic counters.  Diagnostic, no reference counterpart. */
int pf_stream_copy(pf_ctx* ctx, int64_t bytes, int reps, double* seconds);

/* Profiling switch: when on, pf_solve launches each iteration's kernels bracketed by HIP
 * events (no graph) and accumulates per-kernel durations for pf_kernel_times. */
int pf_set_profile(pf_ctx* ctx, int on);

/* Average device time of the two kernels of one PCG iteration (fused operator apply;
 * preconditioner update), each launched `reps` times back to back on the state of
 * the last solve of that precision (the solution is restored).  For benchmarks. */
int pf_time_kernels(pf_ctx* ctx, int precision, int precond, int reps, double* t_apply, double* t_update);

/* Time `reps` back-to-back FP64 operator applies y = A x on the context stream with HIP
 * events; returns the average seconds per apply.  For benchmarks. */
int pf_time_apply(pf_ctx* ctx, int precision, int reps, double* seconds);

/* Picard-loop support ------------------------------------------------------ */

/* pf_solution_prolong for a coarse field that is already in device memory (e.g. the output of
 * pf_net_predict with out = NULL, pf_net_output_device): x = offset + scale * coarse_dev,
 * trilinear, then the Dirichlet values -- the same arithmetic, so the same result bit for
 * bit -- and this entry point allocates nothing: no hipMalloc, hipFree or allocating
 * library call.  No counterpart in the reference; it is for callers that cannot allocate
 * device memory (tight serving loops, code that must not allocate between solves).  Its
 * per-axis tables (4 + 4 bytes per node of each fine axis) live in a fixed 256 KiB device
 * workspace, one per GPU for the life of the process: grids whose three axes have more
 * than 32768 nodes in all are refused (PF_ESHAPE).  Calls on one GPU share that workspace:
 * calls that overlap in time on one GPU (from several threads) race; each call has finished
 * with it when it returns.  coarse_dev must stay valid until the call returns. */
int pf_solution_prolong_dev(pf_ctx* ctx, const float* coarse_dev, const double* cx, int ncx, const double* cy,
                            int ncy, const double* cz, int ncz, double offset, double scale);

/* Set the device solution (N values), e.g. the initial field inittemp (therm2d.cpp:79). */
int pf_solution_set(pf_ctx* ctx, const double* x);

/* Start vector from node values on a coarser rectilinear grid of the same domain
 * (3D): x = offset + scale * coarse, trilinear in physical coordinates, then the
 * Dirichlet values; solve next with x0 = PF_X0_DEVICE.  Used for the learned
 * warm start (the network predicts on its own grid).  No reference counterpart. */
int pf_solution_prolong(pf_ctx* ctx, const float* coarse, const double* cx, int ncx, const double* cy, int ncy,
                        const double* cz, int ncz, double offset, double scale);

/* max |X - X_saved| and max X, then X_saved = X: the loop error and maxT of
 * ThermalFem2DSolver::compute (therm2d.cpp:479-487). */
int pf_solution_change(pf_ctx* ctx, double* max_change, double* max_value);

/* Element means of the device solution (the element temperature T_avg of
 * therm2d.cpp:210 / therm3d.cpp:209), E values to the host. */
int pf_element_means(pf_ctx* ctx, double* out);

/* Temperature-dependent conductivity law per element, k = k0 (300 / T_avg)^alpha
 * (alpha = 0: constant); pf_coeffs_update evaluates it on the device from the current
 * solution, replacing Material::thermk(T) calls of setMatrix (therm3d.cpp:211-213).  An
 * element whose k0_h is NaN keeps the conductivities it has (junction elements, whose
 * conductivity the junction update owns: loadConductivities, electr2d.cpp:331-353). */
int pf_coeffs_powerlaw(pf_ctx* ctx, int equation, const double* k0_h, const double* k0_v, const double* a_h,
                       const double* a_v, int64_t nelem);
int pf_coeffs_update(pf_ctx* ctx);

/* The same law evaluated at the element temperatures of another context's solution on the
 * same grid and device, device to device: the electrical solver's conductivities at the
 * thermal solver's temperatures (ElectricalFem3DSolver::loadConductivities reading
 * inTemperature, electr3d.cpp:324-345, with inTemperature = thermal outTemperature,
 * thermoelectric.py:69-84; the linear interpolation of a same-mesh temperature at element
 * midpoints is the corner mean the law uses). */
int pf_coeffs_update_from(pf_ctx* ctx, const pf_ctx* temperature_src);

/* Element heat source of dst from another context's field on the same grid and device,
 * device to device.  what = PF_POST_JOULE: the Joule heat density of src's last
 * pf_currents (thermal.inHeat = electrical.outHeat, thermoelectric.py:69-84; the element
 * field read at its own element midpoints is the value itself). */
int pf_source_from(pf_ctx* dst, const pf_ctx* src, int what);

/* Transient thermal solve (DynamicThermalFem2DSolver / DynamicThermalFem3DSolver,
 * solvers/thermal/dynamic/femT2d.cpp, femT3d.cpp: the theta method of setMatrix
 * :140-279 and of the time loop of compute :282-331): with the stiffness K of
 * pf_coeffs_set / pf_coeffs_update and the load of pf_source_set / pf_load_set, each step
 * solves A T^{n+1} = B T^n + F with A = theta K + C, B = -(1 - theta) K + C, C the lumped
 * heat-capacity matrix (the reference's default, lumping = true): each node of an element
 * gets c_elem[e], the element's capacity over the time step in the units of K as the
 * reference forms it (3D: cp rho 0.125e-9 dx dy dz / timestep; 2D: cp rho 0.25e-12 w h /
 * timestep / 1e-9, times the mid radius in cylindrical geometry).  theta = methodparam
 * (0 explicit, 0.5 Crank-Nicolson, 1 implicit).  c_elem = NULL returns to the steady
 * operator.  While a capacity is set, pf_solve solves with A and pf_get_coeffs returns K. */
int pf_capacity_set(pf_ctx* ctx, const double* c_elem, int64_t nelem, double theta);

/* Dirichlet nodes in a time step (pf_dynamic_step's held argument).
 * PF_HELD_FIXED: a Dirichlet node keeps its boundary value (the theta method on the
 *   free nodes).
 * PF_HELD_REFERENCE: the reference's update.  Its setMatrix eliminates the conditions from A
 *   and F only (A.applyBC(btemperature, F), femT2d.cpp:243, femT3d.cpp:234) and B is left
 *   whole, so a Dirichlet node's new value is its boundary value plus the (B T^n) of its row
 *   (the free rows use the boundary value).  With realistic 2D capacities that term grows
 *   geometrically; it is offered for comparisons with the reference. */
#define PF_HELD_FIXED 0
#define PF_HELD_REFERENCE 1

/* One time step from the device temperature (the previous step or pf_solution_set): the
 * right-hand side B T^n + F on the device, then pf_solve with opts (x0 forced to
 * PF_X0_DEVICE: the step starts from T^n; PF_NORM_START matches a direct solve per step).
 * The reference factorises A once per compute() and back-substitutes every step
 * (femT2d.cpp:413-432; femT3d.cpp:304-317, whose extra swap of temperatures and X at :306
 * leaves the right-hand side, not the solution, in temperatures: this entry point follows
 * the 2D loop in both dimensions). */
int pf_dynamic_step(pf_ctx* ctx, const pf_solve_opts* opts, int held, pf_solve_stats* stats);

/* Element fields after a solve (the providers' post-processing, on the device):
 * PF_POST_FLUX    heat flux -k grad T (W/m^2; dim values per element; saveHeatFluxes,
 *                 therm2d.cpp:504-540, therm3d.cpp:342-384), with k_h / k_v the element
 *                 conductivities to use (the reference evaluates thermk at the final element
 *                 temperatures with the leaf's height), or NULL: the conductivity law of
 *                 pf_coeffs_powerlaw at the current temperatures, else the conductivities as set;
 * PF_POST_CURRENT current density (kA/cm^2; dim values) and
 * PF_POST_JOULE   Joule heat density (W/m^3) of the last pf_currents (electr2d.cpp:395-464).
 * The field stays on the device for pf_interp_elements; out (E or E * dim values) may be NULL. */
#define PF_POST_FLUX 0
#define PF_POST_CURRENT 1
#define PF_POST_JOULE 2
int pf_postprocess(pf_ctx* ctx, int what, const double* k_h, const double* k_v, double* out);
/* An element field of pf_postprocess at points (npts x dim, um) into out (npts x components):
 * linear between element midpoints, the outermost midpoint's value up to the mesh boundary,
 * fill outside the mesh (interpolation on the element mesh of getCurrentDensities /
 * getHeatDensities / getHeatFluxes, electr2d.cpp:518-562, therm2d.cpp:556-569). */
int pf_interp_elements(pf_ctx* ctx, int what, const double* pts, int64_t npts, double fill, double* out);

/* Current element conductivities (E values each) to the host. */
int pf_get_coeffs(pf_ctx* ctx, double* k_inplane, double* k_vert);

/* Junction elements of the Shockley solver (active regions found by
 * ElectricalFem2DSolver::setupActiveRegions, electr2d.cpp:90-169): for element q its
 * index, ncorner (2 in 2D, 4 in 3D) bottom/top node pairs spanning the whole junction
 * (act.bottom/act.top rows, electr2d.cpp:250-253, electr3d.cpp:253-259), and its
 * junction number; heights[nj] = junction thickness d (um). */
int pf_junctions_set(pf_ctx* ctx, int64_t nje, int ncorner, const int64_t* elems, const int64_t* bottom_nodes,
                     const int64_t* top_nodes, const int* junction, int nj, const double* heights);
/* beta (1/V) and js (A/m^2) per junction element (BetaSolver getBeta/getJs, beta.hpp:50-77). */
int pf_junction_params(pf_ctx* ctx, const double* beta, const double* js, int64_t nje);
/* One junction conductivity update from the device potential (the loopno != 0 block of
 * ElectricalFem2DSolver::setMatrix, electr2d.cpp:243-264; law beta.hpp:43-46).
 * stable = 1: CONVERGENCE_STABLE averaging. */
int pf_junction_update(pf_ctx* ctx, int stable);
/* The inputs of a junction law given by the caller (PythonCondSolver::activeCond of the
 * ActiveCond solvers, electr_python.cpp:171-178, called from electr2d.cpp:247-256): per
 * junction element the junction voltage U (V; mean over the element's columns of the
 * potential at the active region's top minus its bottom) and the vertical current density
 * jy = 0.1 sigma_v U / height (kA/cm^2, signed), from the device potential and the current
 * conductivities.  Either pointer may be NULL. */
int pf_junction_state(pf_ctx* ctx, double* U, double* jy, int64_t nje);
/* Set the junction elements' conductivities (S/m, in-plane and vertical) from such a law:
 * stable = 1 averages with the current values (CONVERGENCE_STABLE), then a vertical value
 * that is NaN or below 1e-16 in magnitude becomes 1e-16 (electr2d.cpp:257-262). */
int pf_junction_cond_set(pf_ctx* ctx, const double* cond_h, const double* cond_v, int64_t nje, int stable);

/* Per-element flags: select = elements whose |j| enters the loop error (junction
 * elements, electr2d.cpp:410; NULL = all), noheat = elements with zero Joule heat
 * (EMPTY material or 'noheat' role, electr2d.cpp:454). */
int pf_element_flags_set(pf_ctx* ctx, const unsigned char* select, const unsigned char* noheat);

/* Element current densities (kA/cm^2) and Joule heat (W/m^3) from the device potential;
 * returns max |j - j_prev|^2 and max |j|^2 over the selected elements (the loop error
 * of electr2d.cpp:399-422) and keeps j as j_prev. */
int pf_currents(pf_ctx* ctx, double* max_dj2, double* max_j2);
int pf_get_currents(pf_ctx* ctx, double* cur, double* heat);

/* Batched uneven 2D Shockley problems --------------------------------------
 * Many independent ElectricalFem2DSolver::compute runs (electr2d.cpp:363-440, junction
 * law beta.hpp:43-46) -- e.g. voltage/temperature sweeps over device variants -- in one
 * persistent kernel launch replayed from a hipGraph.  Each workgroup takes instances
 * from a largest-first work queue and runs the whole Picard loop on the device: junction
 * update, line- or multigrid-preconditioned CG (warm start), currents and loop error.  The reference
 * runs such sweeps as separate processes (gui/plugins/launcher_batch.py:16). */
typedef struct pf_batch pf_batch;

typedef struct {
    int maxloops;   /* Picard loops (compute(loops)), default 30 */
    int maxit;      /* CG iterations per loop */
    double maxerr;  /* loop error limit in % (ElectricalFem2DSolver::maxerr, default 0.05) */
    double rtol;    /* CG of each Picard loop: ||B - A V|| <= rtol max(||B - A V_start||, 1e-4 ||B - A V_init||),
                       V_start the previous loop's potential (the loop's own start residual reduced by rtol, as
                       a direct solve of every loop would), V_init = Dirichlet values, 0 elsewhere.  The true
                       residual is recomputed with compensated (double-double) sums and CG restarted once from
                       it when it has drifted above that level (a loop that still misses it -- junction systems
                       with conductivities spanning ~1e14 -- has the potential FP64 CG reaches; `rel` and `bwd`
                       report it) */
    int stable;     /* CONVERGENCE_STABLE */
    int block;      /* threads per workgroup 256/512/1024; 0 = autotuned or 1024 */
    int nwg;        /* persistent workgroups; 0 = autotuned or one per CU */
    int precond;    /* 3 (or 0, the default) = multigrid V-cycle (axis 0 coarsened per instance,
                       damped line smoother, omega 0.65, 2 coarsest sweeps); 2 = line preconditioner */
    int loop_budget;/* Picard loops one kernel launch runs per instance (0: all of them in one launch).
                       The solve relaunches the kernel until every instance has finished, so a
                       launch returns after a bounded amount of work whatever the batch */
    /* called between launches (may be NULL): launches so far, instances finished, total,
       seconds since the first launch, progress_user */
    void (*progress)(int launches, int finished, int total, double seconds, void* user);
    void* progress_user;
} pf_batch_opts;

typedef struct {
    int loops, iters, status; /* status 0 ok, 1 CG maxit, 2 Picard maxloops, <0 breakdown */
    int restarts;    /* CG restarts from a recomputed (compensated) true residual, summed over loops */
    double err;      /* last loop error (%) */
    double maxj;     /* max |j| at the junction (kA/cm^2) */
    double rel;      /* last loop: final true residual over its reference residual (see rtol) */
    double current;  /* total current through the middle junction row (mA; 2D Cartesian per um of length) */
    double bwd;      /* last loop: componentwise backward error max_i |r_i| / (|A| |V|)_i of its potential (a
                        backward-stable direct solve reaches a small multiple of eps = 2.2e-16) */
} pf_batch_stats;

int pf_batch_create(pf_batch** out, int device);
void pf_batch_destroy(pf_batch* b);
const char* pf_batch_last_error(const pf_batch* b);
/* ninst instances; per instance node counts n0 (axis 0 = r) and n1 (axis 1 = z); node
 * axes, element conductivities (material values, natural order; junction elements hold
 * the start value (0, start_cond)), Dirichlet (local node, value) lists with ndir
 * entries each, junction element rows [jlo, jhi) (jlo == jhi: no junction), beta, js. */
int pf_batch_set(pf_batch* b, int ninst, int cylindrical, const int32_t* n0, const int32_t* n1, const double* ax0,
                 const double* ax1, const double* sh, const double* sv, const int64_t* ndir, const int64_t* dnodes,
                 const double* dvals, const int32_t* jlo, const int32_t* jhi, const double* beta, const double* js);
/* Times each (block, workgroups) schedule on the batch and keeps the fastest. */
int pf_batch_autotune(pf_batch* b, const pf_batch_opts* o, int* best_block, int* best_nwg, double* best_seconds);
/* Solve every instance from V = Dirichlet values / 0; V_out (concatenated, may be NULL),
 * stats[ninst] (may be NULL), seconds = device time of all launches (the start launch and
 * every launch of the persistent kernel, with the host's progress reads between them).
 * The launches are replayed from hipGraphs (plain launches under a kernel-tracing
 * profiler, as for pf_solve). */
int pf_batch_solve(pf_batch* b, const pf_batch_opts* o, double* V_out, pf_batch_stats* stats, double* seconds);
/* Element conductivities (in-plane, vertical; concatenated, natural order per instance) the
 * last loop of the last pf_batch_solve used: the material values, and at the junction the
 * values of the last junction update (saveConductivities, electr2d.cpp:355-361).  Either
 * pointer may be NULL. */
int pf_batch_conductivities(pf_batch* b, double* sh, double* sv);

#ifdef __cplusplus
}
#endif
#endif
