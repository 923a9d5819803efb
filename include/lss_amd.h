/*
 * lss_amd.h -- C-ABI of the MI355X Heston/SABR ADI library (liblss_amd.so).
 *
 * Plain C types only: sizes, enums, host or device pointers and a hipStream_t passed as void*.
 * The C++ drop-in headers (include/lss/...) and the Python bindings (linear_system_solvers_amd)
 * both sit on top of these entry points.  Enum values are those of the reference's
 * common/lss_enumerations.hpp:15-129.
 *
 * Status: every entry point returns LSS_OK (0) or a negative lss_status; no C++ exception
 * crosses this boundary.  lss_last_error() returns a message for the calling thread's last error.
 */
#ifndef LSS_AMD_H
#define LSS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lss_status {
    LSS_OK = 0,
    LSS_ERR_INVALID = -1,     /* bad argument, size or combination                  */
    LSS_ERR_UNSUPPORTED = -2, /* valid in the reference API but not offered here    */
    LSS_ERR_DEVICE = -3,      /* HIP runtime error                                  */
    LSS_ERR_ALLOC = -4        /* host or device allocation failed                   */
} lss_status;

enum { LSS_FP64 = 0, LSS_FP32 = 1 };
/* splitting_method_enum; LSS_EXPLICIT_EULER selects the explicit solver instead
 * (explicit_solvers::general_heston_equation with explicit_pde_schemes_enum::Euler,
 * lss_2d_general_heston_equation.hpp:430-503, lss_heston_euler_solver_scheme.hpp:46-92): per step the
 * Euler update of the interior (lss_heston_euler_solver_method.hpp:26-52), then the explicit boundary
 * solves A and B shared with the ADI schemes.  A plan steps either explicit Euler problems only or
 * ADI problems only; explicit plans have no adjoint. */
enum { LSS_DOUGLAS_RACHFORD = 0, LSS_CRAIG_SNEYD = 1, LSS_MODIFIED_CRAIG_SNEYD = 2, LSS_HUNDSDORFER_VERWER = 3,
       LSS_EXPLICIT_EULER = 4 };
/* implicit_pde_schemes_enum */
enum { LSS_EULER = 0, LSS_THETA_30 = 1, LSS_CRANK_NICOLSON = 2, LSS_THETA_80 = 3 };
/* tridiagonal_method_enum: selects the line solver's operation order */
enum { LSS_CUDA_SOLVER = 0, LSS_DOUBLE_SWEEP = 1, LSS_SOR_SOLVER = 2, LSS_THOMAS_LU = 3 };
/* traverse_direction_enum */
enum { LSS_FORWARD = 0, LSS_BACKWARD = 1 };
/* grid_enum */
enum { LSS_GRID_UNIFORM = 0, LSS_GRID_NONUNIFORM = 1 };
/* coefficient representation of a plan */
enum {
    LSS_COEFF_FIELDS = 0,    /* per point M, M~, P, P~, Z, W, C + D, E, F on the v = 0 row        */
    LSS_COEFF_SEPARABLE = 1  /* q(i,j) = X0q[i] Y0q[j] + X1q[i] Y1q[j] for q = a, b, c, d, e, f     */
};
/* line solver of the implicit sweeps */
enum {
    LSS_LINE_EXACT = 0,       /* one lane per line in the reference's Thomas / double-sweep order   */
    LSS_LINE_PARTITIONED = 1  /* register-resident chunks + cyclic reduction of the chunk interfaces */
};
/* batched tridiagonal layouts */
enum { LSS_LAYOUT_CONTIGUOUS = 0, LSS_LAYOUT_STRIDED = 1 };

/*
 * Sampled description of one ADI problem, in the precision `fp` (arrays are double* when fp is
 * LSS_FP64 and float* when LSS_FP32).  It carries what general_heston_equation::solve evaluates
 * through its closures (pde_solvers/two_dimensional/heat_type/lss_2d_general_heston_equation.hpp:184-307):
 *   coeff      coeff_sets sets; FIELDS set = 7 nx*ny fields (M, M~, P, P~, Z, W, C at (t, zeta_i, eta_j),
 *              coefficients.hpp:89-95) then D, E, F on the v = 0 row (3 nx); SEPARABLE set = 12 nx
 *              vectors X[2q+r] then 12 ny vectors Y[2q+r] with transformed q = a..f
 *              (lss_heat_data_transform.hpp:174-214).  coeff_sets is 1 (time independent) or steps.
 *   bnd        bnd_sets sets of nx + 2 ny values: v_upper(t, X(zeta_i)), h_lower(t, Y(eta_j)),
 *              h_upper(t, Y(eta_j)) * X'(1) (lss_heston_boundary_transform.hpp:48-62).
 *   idx1/idx2  index_of maps (lss_grid_config.hpp:110-118).
 *   theta..h1  scheme scalars, already rounded to fp (implicit_kernel.hpp:978-995, coefficients.hpp:62-73).
 *   src        heat source: src_sets == 0 (none, src NULL), 1 (time independent) or steps sets of
 *              2 nx*ny values, F(t, X(zeta_i), Y(eta_j)) at loop step s's time t and at t_next
 *              (t - k on a Backward loop, t + k Forward), row-major; stage 1 adds
 *              theta rho F(t_next) + (1 - theta) rho F(t) (rhs_intermed_1_source,
 *              lss_heat_douglas_rachford_method.hpp:70-104 and the same function of CS/MCS/HV, which the
 *              reference defines but does not call: its loop returns without stepping when a source is
 *              set, implicit_kernel.hpp:1032-1044).  ADI plans only.
 * FIELDS with LSS_LINE_EXACT reproduces the reference's host result bit for bit.  The partitioned
 * solver needs (nx - 2) and (ny - 2) <= 64 * 16.
 * All pointers are host pointers; the plan copies them to device memory.
 */
typedef struct lss_adi_desc {
    int32_t fp;
    int32_t splitting;
    int32_t tridiagonal;     /* LSS_THOMAS_LU or LSS_DOUBLE_SWEEP operation order; LSS_CUDA_SOLVER = Thomas */
    int32_t coeff_mode;
    int32_t line_solver;     /* LSS_LINE_EXACT | LSS_LINE_PARTITIONED; equal for all problems of a plan */
    int32_t pad;
    int64_t nx, ny, steps;
    double theta, weighting;
    double alpha, beta, gamma, delta, ni, rho, h1;
    const uint32_t *idx1, *idx2;
    int64_t coeff_sets;
    const void *coeff;
    int64_t bnd_sets;
    const void *bnd;
    int64_t src_sets;
    const void *src;
} lss_adi_desc;

typedef struct lss_adi_plan lss_adi_plan;

const char *lss_last_error(void);
int32_t lss_version(void);

/*
 * Plans.  A plan holds `count` independent problems (count == 1: the single-problem stepper that
 * general_heston_equation::solve's device branch uses, heston_equation.hpp:221-233; count > 1: a
 * batched book stepping all problems in the same launches).  All problems share fp.
 */
int lss_adi_plan_create(const lss_adi_desc *descs, int64_t count, lss_adi_plan **plan);
int lss_adi_plan_destroy(lss_adi_plan *plan);

/*
 * Time loop (implicit_heston_time_loop::run, implicit_kernel.hpp:150-207) for every problem of
 * the plan.  u_in[p] holds problem p's initial (terminal) surface, u_out[p] receives the final
 * surface; nx*ny row-major each.  _dev: device pointers, work enqueued on `stream` (a hipStream_t,
 * NULL = default stream), returns without synchronising.  _host: host pointers, synchronous.
 * use_graph != 0 replays the step launches from a hipGraph captured on first use.
 */
int lss_adi_run_dev(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, void *stream, int32_t use_graph);
/*
 * Autotuner (no counterpart in the reference): times the tile/lane mappings of the partitioned sweeps
 * (S-lines per x-sweep workgroup: 16, 8 or 4; v-lines per v-sweep workgroup: 4 or 16; tiles in launch
 * order or XCD-contiguous order) on the first steps of the loop and keeps the fastest; the mappings
 * give identical results.  Runs by itself before a plan's first solve; `again` != 0 re-times (and
 * drops a captured graph).  Overwrites the plan's state buffers, synchronous with respect to
 * `stream`.  Returns the chosen mapping.
 */
int lss_adi_autotune(lss_adi_plan *plan, void *stream, int32_t again, int32_t *s_lines_per_group,
                     int32_t *v_lines_per_group, int32_t *xcd_order);
/* Sets the mapping instead of timing it (16, 8 or 4 S-lines, 4 or 16 v-lines per workgroup, 0 or 1). */
int lss_adi_set_mapping(lss_adi_plan *plan, int32_t s_lines_per_group, int32_t v_lines_per_group,
                        int32_t xcd_order);
/* (no counterpart in the reference) How lss_adi_backward schedules its recomputation: 1 always with
 * a second segment store on a second stream, 0 never, -1 (the default) when the second store takes
 * at most 2 GiB.  Takes effect at the next lss_adi_forward. */
int lss_adi_set_adjoint_overlap(lss_adi_plan *plan, int32_t mode);
/* Kernel configuration of a plan (LSS_KERNEL_CONFIG_LEN values): line solver, S-lines and v-lines per
 * workgroup, XCD order, compile-time factor pattern of the coefficients (0 none, 1 Heston, 2 SABR with
 * r = 0), chunk bounds of S-lines and v-lines, whether the autotuner has run, the compile-time pattern
 * of the adjoint kernels (factor and derivative-term tables; -1 before the adjoint's first use), the
 * S-lines per adjoint S-sweep workgroup (0 before first use) and whether the backward pass overlaps
 * the recomputation of the next segment with the adjoint of the current one (two segment storage
 * sets, see lss_adi_set_adjoint_overlap; -1 before first use), and the number of
 * concurrent groups a graph replay of lss_adi_run_dev runs the problems in (2 for plans of two or
 * more problems, else 1), and whether lss_adi_run_dev / lss_adi_run_host step an explicit Euler plan in
 * one launch with the time levels resident in LDS (1: every problem has at most 1536 points, nx >= 4,
 * ny <= 256, one coefficient set and one boundary set; lss_adi_run_levels always steps launch by
 * launch). */
enum { LSS_KERNEL_CONFIG_LEN = 13 };
int lss_plan_kernel_config(const lss_adi_plan *plan, int32_t *config);
int lss_adi_run_host(lss_adi_plan *plan, const void *const *u_in, void *const *u_out);
/*
 * Every time level of the loop (solve(container_3d&), lss_2d_general_heston_equation.hpp:180, 309-313,
 * whose body the reference leaves empty; the layering follows the reference's 1-D run_with_stepping,
 * lss_1d_general_heat_equation_implicit_kernel.hpp:163-257): levels[p] receives steps + 1 surfaces of
 * nx*ny values, row s being the surface after s steps of the loop (row 0 = u_in[p]).  host = 0: device
 * pointers, enqueued on `stream`; host != 0: host pointers, synchronous.  No graph replay.
 */
int lss_adi_run_levels(lss_adi_plan *plan, const void *const *u_in, void *const *levels, void *stream,
                       int32_t host);

/*
 * Reference-model convenience: sample the Heston or SABR data of the reference's tests
 * (unit_tests/common/lss_print_t.hpp:2156-2186 Heston; :2039-2050 SABR) on the reference's grid and
 * build a descriptor-equivalent problem inside the library.  One lss_model per problem.
 */
enum { LSS_MODEL_HESTON = 0, LSS_MODEL_SABR = 1 };
typedef struct lss_model {
    int32_t kind;              /* LSS_MODEL_HESTON | LSS_MODEL_SABR */
    int32_t pad;
    /* Heston: a = 0.5 v S^2, b = 0.5 sig^2 v, c = rho sig v S, d = r S, e = kappa (theta_v - v), f = -r
     * SABR:   D = exp(-r (T - t)); a = 0.5 al^2 F^(2 beta) D^(2(1-beta)), b = 0.5 sig^2 al^2,
     *         c = rho sig al^2 F^beta D^(1-beta), d = r F, e = 0, f = -r                        */
    double strike, rate, sig_sig, sig_kappa, sig_theta, rho, beta, maturity;
    double bnd_low;            /* Dirichlet value at x_lo (reference tests: 0)  */
    double bnd_high;           /* Neumann value at x_hi (reference tests: -1)   */
    /* pde_discretization_config_2d + grid_config_hints_2d + splitting/solver config */
    double x_lo, x_hi, y_lo, y_hi, t_lo, t_hi;
    int64_t nx, ny, nt;
    double accumulation_point, alpha_scale, beta_scale;
    int32_t grid;              /* LSS_GRID_UNIFORM | LSS_GRID_NONUNIFORM */
    int32_t splitting;
    double weighting;
    int32_t implicit_scheme;
    int32_t tridiagonal;
    int32_t direction;
    int32_t coeff_mode;        /* LSS_COEFF_FIELDS (reference rounding) or LSS_COEFF_SEPARABLE */
    int32_t line_solver;       /* LSS_LINE_EXACT | LSS_LINE_PARTITIONED */
} lss_model;

int lss_model_plan_create(const lss_model *models, int64_t count, int32_t fp, lss_adi_plan **plan);
/* sampled initial (terminal) surface of problem p of a model plan, host pointer, nx*ny */
int lss_model_initial(const lss_adi_plan *plan, int64_t p, void *u0_host);
/* physical grid of problem p of a model plan: S (or F) values X(zeta_i), nx, and v (or alpha)
 * values Y(eta_j), ny (lss_grid_transform_config.hpp:118-149), host pointers in the plan's precision */
int lss_model_grid(const lss_adi_plan *plan, int64_t p, void *xs_host, void *ys_host);
/* sizes of problem p */
int lss_plan_problem_shape(const lss_adi_plan *plan, int64_t p, int64_t *nx, int64_t *ny, int64_t *steps);

/*
 * Batched plain tridiagonal systems (replaces one thomas_lu_solver::solve per line,
 * sparse_solvers/tridiagonal/thomas_lu_solver/lss_thomas_lu_solver.hpp:125-191, with Neumann(0) ends
 * and a[0] = c[n-1] = 0, i.e. the plain algorithm).  B systems of length n; layout CONTIGUOUS
 * (x[k*n+i]) or STRIDED (x[i*B+k]); device pointers; enqueued on `stream`.  Systems of up to 1024
 * unknowns (contiguous fp64: 512) are solved by the register-resident partitioned elimination with
 * no scratch; longer ones one lane per system, with B*n values of scratch allocated on `stream`
 * (stream-ordered, hipMallocAsync) and freed on it after the solve, so that calls on different streams
 * never share scratch.
 */
int lss_tridiag_batched(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                        const void *c, const void *d, void *x, void *stream);

/*
 * The same batched solve with end conditions folded into the first and last rows as the reference's
 * thomas_lu_solver folds its boundary pair (lss_thomas_lu_boundary.hpp:114-205, with
 * lss_dirichlet/neumann/robin_boundary.hpp conventions), the same for every system:
 *   LSS_BC_DIRICHLET  the end value is val (x_0 or x_{n-1} = val; the system is solved for the rest);
 *   LSS_BC_NEUMANN    u' + val = 0: the ghost point x_{-1} = x_1 + 2 h val (lower) or
 *                     x_n = x_{n-2} - 2 h val (upper) is folded into the end row;
 *   LSS_BC_ROBIN      u' + lin u + val = 0: ghost x_{-1} = x_1 + 2 h (lin x_0 + val), x_n likewise;
 * h = 1 / (n - 1).  Replaces thomas_lu_solver::solve(boundary_pair, solution) per system.
 */
enum { LSS_BC_DIRICHLET = 0, LSS_BC_NEUMANN = 1, LSS_BC_ROBIN = 2 };
typedef struct {
    int32_t type; /* LSS_BC_* */
    double lin;   /* Robin only */
    double val;
} lss_line_bc;
int lss_tridiag_batched_bc(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                           const void *c, const void *d, const lss_line_bc *lower, const lss_line_bc *upper, void *x,
                           void *stream);

/*
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the solve of
 * lss_tridiag_batched (lower == upper == NULL) or lss_tridiag_batched_bc (both non-NULL) that calls
 * no hipMalloc/hipFree or allocating routine.  The one-lane-per-system kernel (systems of more than
 * 1024 unknowns, or contiguous fp64 systems of more than 512) takes its B*n values of scratch from a
 * static per-device array of LSS_TRIDIAG_SCRATCH_BYTES, which each device holds until the process
 * ends; larger B*n*sizeof(fp) is rejected with LSS_ERR_INVALID.  Every call on a device shares that
 * array: calls that overlap on one device (different streams or threads) race and give wrong results.
 * The caller passes inputs and outputs as for lss_tridiag_batched.
 */
#define LSS_TRIDIAG_SCRATCH_BYTES ((int64_t)64 << 20)
int lss_tridiag_batched_noalloc(int32_t fp, int32_t layout, int64_t B, int64_t n, const void *a, const void *b,
                                const void *c, const void *d, const lss_line_bc *lower, const lss_line_bc *upper,
                                void *x, void *stream);

/*
 * Same time loop as lss_adi_run_dev (device pointers, no graph) with a hipEvent pair recorded on
 * `stream` around every launch; synchronous.  For each kernel kind k < LSS_K_COUNT it returns the
 * summed launch time ms[k], the summed algorithmic HBM bytes bytes[k] (compulsory reads and writes
 * of the 2-D state arrays of the problems active in each launch) and the number of launches.
 */
enum {
    LSS_K_BND_A = 0, LSS_K_XSWEEP1 = 1, LSS_K_VSWEEP1 = 2, LSS_K_HV_Y3 = 3, LSS_K_XSWEEP2 = 4, LSS_K_VSWEEP2 = 5,
    LSS_K_BND_B = 6, LSS_K_EULER = 7, LSS_K_COUNT = 8
};
int lss_adi_run_profiled(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, void *stream, double *ms,
                         double *bytes, int64_t *launches);

/*
 * Reverse mode (no counterpart in the reference, whose solvers have no derivatives): the adjoint of
 * the time loop, shaped as an autograd-style forward/backward pair.
 *   lss_adi_forward   the time loop of lss_adi_run_dev (device pointers, enqueued on `stream`) that also
 *                     keeps every ckpt_every-th time level (ckpt_every even, >= 2) in plan memory.
 *   lss_adi_backward  given lam_out[p] = dJ/du_out[p] (device, nx*ny), recomputes each segment of
 *                     ckpt_every steps from its checkpoint and runs the adjoint steps backwards.  Writes
 *                     lam_in[p] = dJ/du_in[p] (device; a NULL entry or array skips it) and, on the host,
 *                     dparams[p * LSS_MAX_PARAMS + k] = dJ/dparam_k of problem p.  Synchronous.
 *                     Needs nx - 2 <= 1024 and ny - 2 <= 1024.  With a second segment store
 *                     (ckpt_every + 1 levels and the stage outputs of ckpt_every steps, over all
 *                     problems; see lss_adi_set_adjoint_overlap) the recomputation of the next
 *                     segment runs on a stream of the plan's own (ordered after the work already
 *                     queued on `stream`) while the adjoint of the current one runs on `stream`;
 *                     every result is ready when the call returns.
 * Parameters enter through the transformed coefficients: a derivative term says
 *   d q(i, j) / d param_k = X_t[i] * Y_t[j]   for q one of a..f (LSS_Q_*),
 * i.e. the derivative of heat_data_transform's coefficient at grid point (i, j)
 * (lss_heat_data_transform.hpp:174-214).  Model plans carry their terms (Heston: kappa, theta_v, sigma,
 * rho, r; SABR: nu, rho, beta and lam, the scale of the alpha axis about y_lo: the grid, and with it a
 * price read at a fixed eta, stretched by lam, so that dPrice/dalpha0 = dJ/dlam / alpha0 when y_lo = 0
 * and y_hi is a multiple of alpha0); other plans set them with lss_adi_set_param_terms: term t has
 * parameter term_param[t] and coefficient term_coeff[t]; factors holds `sets` (1 or steps) blocks of
 * nterm * (nx X-values then ny Y-values) in the plan's precision.
 */
enum { LSS_MAX_PARAMS = 6, LSS_MAX_TERMS = 12 };
enum { LSS_Q_A = 0, LSS_Q_B = 1, LSS_Q_C = 2, LSS_Q_D = 3, LSS_Q_E = 4, LSS_Q_F = 5 };
int lss_adi_set_param_terms(lss_adi_plan *plan, int64_t p, int32_t nparam, int32_t nterm, const int32_t *term_param,
                            const int32_t *term_coeff, int64_t sets, const void *factors);
int lss_plan_param_count(const lss_adi_plan *plan, int64_t p, int32_t *nparam);
int lss_adi_forward(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, int32_t ckpt_every, void *stream);
int lss_adi_backward(lss_adi_plan *plan, const void *const *lam_out, void *const *lam_in, double *dparams,
                     void *stream);
/* The same pair on host pointers (staged through device memory; synchronous). */
int lss_adi_forward_host(lss_adi_plan *plan, const void *const *u_in, void *const *u_out, int32_t ckpt_every);
int lss_adi_backward_host(lss_adi_plan *plan, const void *const *lam_out, void *const *lam_in, double *dparams);

#ifdef __cplusplus
}
#endif
#endif
