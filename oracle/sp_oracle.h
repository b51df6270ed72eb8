/* CPU oracle for the ePlace mGP/cGP evaluation and iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (superplace_amd) never
 * links, loads or falls back to it.
 *
 * It is a from-scratch restatement of the reference's CPU algorithm, in two
 * arithmetic modes chosen per call:
 *   precision 64 : every quantity in double ("the truth" the GPU is compared to);
 *   precision 32 : the reference's own arithmetic, FP32 values with FP64
 *                  fixed-chunk reductions (used as the CPU baseline and for
 *                  bit-exact checks of the fixed-point density scatter).
 * Each function names the reference file:line it follows.  Functions whose
 * parity is not pinned by a golden vector, closed form or brute force say
 * "parity unpinned" below; at present every exported function is pinned by
 * tests/test_oracle_*.py.
 */
#ifndef SP_ORACLE_H
#define SP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node flags, src/db/place_db.h:17-22. */
enum { ORA_F_MACRO = 1, ORA_F_FIXED = 2, ORA_F_FILLER = 4, ORA_F_NI = 8 };

/* Pure SoA + CSR placement database, src/db/place_db.h:27-71.
 * Node order [movable | fixed | filler]; coordinates are node centres.
 * node2pin_start has num_nodes+num_fillers+1 entries (fillers have no pins,
 * src/gp/filler.cpp:129-130). */
typedef struct ora_design {
    int32_t num_movable, num_nodes, num_fillers, num_pins, num_nets, num_rows;
    const float* node_w;
    const float* node_h;
    const uint8_t* node_flags;
    const float* pin_offset_x;
    const float* pin_offset_y;
    const int32_t* pin2node;
    const int32_t* pin2net;
    const int32_t* net2pin_start;
    const int32_t* flat_net2pin;
    const int32_t* node2pin_start;
    const int32_t* flat_node2pin;
    const float* row_ly;     /* site rows: lower y, height, lower x, upper x */
    const float* row_height;
    const float* row_lx;
    const float* row_hx;
    float core_lx, core_ly, core_hx, core_hy;
} ora_design;

/* ---- bin grid (src/gp/bin_grid.cpp) --------------------------------------- */
typedef struct ora_grid ora_grid;

/* BinGrid::initialize, bin_grid.cpp:172-297 (dimension rule, steps, static
 * terminal and base density, tau denominator). */
ora_grid* ora_grid_create(const ora_design* d, const float* x, const float* y,
                          float target_density, int bin_dim_override, int precision);
void ora_grid_destroy(ora_grid* g);
/* out[0..5] = dim, step_x, step_y, origin_x, origin_y, movable_area_scaled */
void ora_grid_info(const ora_grid* g, double* out6);
/* terminal / base static density grids (dim*dim, row-major, y fast). */
void ora_grid_static(const ora_grid* g, double* terminal, double* base);

/* BinGrid::accumulate, bin_grid.cpp:299-359.  Outputs dim*dim grids. */
void ora_accumulate(const ora_grid* g, const ora_design* d, const float* x, const float* y,
                    double* node_density, double* filler_density);
/* Fixed-point variant: each bin-overlap contribution is the reference's FP32
 * value (footprintOf/overlapArea/chargeWeight in float) rounded to nearest
 * multiple of 2^-frac_bits and summed as int64. */
void ora_accumulate_fixed(const ora_grid* g, const ora_design* d, const float* x,
                          const float* y, int frac_bits, int64_t* node_q, int64_t* filler_q);
/* BinGrid::overflow, bin_grid.cpp:373-393. */
double ora_overflow(const ora_grid* g, const double* node_density);
/* BinGrid::solveField, bin_grid.cpp:361-371: rho = sum of four grids / binArea. */
void ora_density_rho(const ora_grid* g, const double* node_density,
                     const double* filler_density, double* rho);
/* BinGrid::gatherForce, bin_grid.cpp:395-426.  force has 2*(num_nodes+num_fillers). */
void ora_gather(const ora_grid* g, const ora_design* d, const float* x, const float* y,
                const double* field_x, const double* field_y, double* force);
/* BinGrid::totalCharge, bin_grid.cpp:428-446. */
double ora_total_charge(const ora_grid* g, const ora_design* d, const float* x, const float* y);

/* ---- Poisson (src/numeric/cpu/cpu_poisson.cpp:149-207) --------------------- */
/* Direct-summation DCT with FFTW's REDFT10 / REDFT01 / RODFT01 conventions
 * (cpu_poisson.cpp:14-36).  rho, phi, ex, ey are nx*ny, row-major, y fast. */
void ora_poisson(int nx, int ny, double step_x, double step_y, const double* rho,
                 double* phi, double* ex, double* ey, int precision);
/* The three 1-D transforms themselves (length n, unnormalised, FFTW's
 * definitions); kind 0 = REDFT10, 1 = REDFT01, 2 = RODFT01. */
void ora_r2r_1d(int kind, int n, const double* in, double* out);
/* The same transforms through one complex FFT of length 2n (n a power of
 * two): the O(n log n) path of the CPU baseline. */
void ora_r2r_1d_fast(int kind, int n, const double* in, double* out);
/* on != 0: ora_poisson and the stage loop use the FFT path for power-of-two
 * grids (the CPU baseline); 0 (default): direct summation. */
void ora_set_fast_poisson(int on);

/* ---- wirelength ------------------------------------------------------------ */
/* computeGamma, src/gp/wa_wirelength.cpp:18-29 (FP32 as in the reference). */
float ora_compute_gamma(float tau, float bin_step);
/* WaWirelength::compute, wa_wirelength.cpp:92-226.  grad has 2*num_movable. */
double ora_wa(const ora_design* d, const float* x, const float* y, float gamma,
              int ignore_net_degree, double* grad, int precision);
/* computeHPWL, src/db/hpwl.cpp:13-39. */
double ora_hpwl(const ora_design* d, const float* x, const float* y, int precision);

/* ---- schedules (src/gp/lambda_schedule.h:11-25) ----------------------------- */
float ora_lambda_multiplier(double delta_hpwl, double reference);
/* hpwlReference: the lambda schedule's HPWL scale; returns 1 where the
 * reference throws (non-positive or non-finite). */
int ora_hpwl_reference(int relative, float delta_ref, double delta_relative, double initial_hpwl,
                       int filler_only, double* out);

/* ---- fillers (src/gp/filler.cpp:47-139) ------------------------------------- */
/* out: count, width, height, whitespace area, node area scaled, filler area */
void ora_filler_plan(const ora_design* d, const float* x, const float* y, float target_density,
                     float row_height, double total_row_area, double* out6);

/* ---- Nesterov (src/numeric/cpu/nesterov.cpp) on a separable quadratic -------- */
/* f(x) = 0.5 sum c_i (x_i - t_i)^2.  Runs `steps` steps from pos (in/out) with
 * optional upper clamp hi (if clamp_hi is finite).  Returns the objective; evals out. */
double ora_nesterov_quadratic(int n, const float* coef, const float* target, float* pos,
                              int steps, int use_bb, int backtrack, float init_step,
                              float clamp_hi, long* evals);

/* ---- the full ePlace stage loop (src/gp/eplace.cpp:223-526) ----------------- */
typedef struct ora_gp_config {
    float target_density;        /* config.h:77 */
    float target_overflow;       /* config.h:78 */
    float cgp_target_overflow;   /* config.h:79 */
    int32_t gp_max_iter;         /* config.h:83 */
    int32_t ignore_net_degree;   /* config.h:84 */
    int32_t gp_bin_dim;          /* config.h:86 */
    float delta_hpwl_ref;        /* config.h:101 */
    int32_t lambda_relative;     /* lambda_schedule == "relative" */
    double delta_hpwl_relative;  /* config.h:105 */
    int32_t use_bb;              /* config.h:107 */
    int32_t gp_backtrack;        /* config.h:109 */
    int32_t filler_only_iters;   /* config.h:111 */
    int32_t gp_stagnation_window;   /* config.h:116 */
    float gp_stagnation_min_gain;   /* config.h:118 */
    int32_t stage;               /* 0 = mGP, 1 = FILLERONLY, 2 = cGP (eplace.h:24-28) */
    int32_t precision;           /* 32 or 64 */
    int32_t max_backtrack;       /* optimizer.h:223 */
} ora_gp_config;

/* Per-iteration metrics row, src/util/metrics.h:12-28. */
typedef struct ora_iter_metrics {
    int32_t iter;
    int32_t evaluations;
    double hpwl, overflow, lambda, gamma, step_size, grad_norm_wl, grad_norm_den;
    double eval_lambda, eval_gamma, delta_hpwl, lambda_mu, elapsed_ms;
} ora_iter_metrics;

typedef struct ora_gp_result {
    int32_t iterations, converged, stagnated, rollback_iter;
    int64_t evaluations;
    double hpwl, overflow;
    double initial_hpwl, initial_overflow, initial_lambda;
    double seconds_loop;   /* wall time of the iteration loop */
} ora_gp_result;

/* Runs one stage over node positions x/y (all nodes, in/out: delivered
 * positions written back for optimised nodes).  metrics may be NULL; at most
 * max_rows rows are written. */
int ora_eplace_run(const ora_design* d, float* x, float* y, const ora_gp_config* cfg,
                   ora_gp_result* res, ora_iter_metrics* metrics, int max_rows);

/* ---- quadratic initial placement (src/init/quadratic_placer.cpp,
 * src/init/b2b_model.h) ------------------------------------------------ */

/* computeNetBoundPins (src/db/hpwl.cpp:70-102): per net the first pin (in
 * net order) holding the smallest / largest x and y, pins at node centre +
 * offset in float; -1 for empty nets. */
void ora_bound_pins(const ora_design* d, const float* x, const float* y, int32_t* min_x,
                    int32_t* max_x, int32_t* min_y, int32_t* max_y);

/* The Bound2Bound system of one axis (axis 0 = x, 1 = y) over the movable
 * nodes (assemble, quadratic_placer.cpp:43-113, pairs of forEachB2BPair,
 * b2b_model.h:56-82): nets of degree 2 .. ignore_deg, weight
 * net_weight / (deg-1) / max(|dpos|, max(1e-6, min_dist)) in float.
 * Row i lists, for each pin p of node i (node->pin order) and each partner q
 * of p (net order; partners of a bound pin: every pin on another node; of
 * any other pin: the bound pins on other nodes) with node(q) movable, the
 * entry (node(q), -w).  diag and rhs are summed in double in the same order;
 * a zero diagonal becomes 1e-6 with rhs += 1e-6 * position.  net_weight may
 * be NULL (all 1).  Returns the entry count; row_start (num_movable + 1),
 * col and val may be NULL to ask for it. */
int64_t ora_b2b_system(const ora_design* d, const float* x, const float* y,
                       const float* net_weight, int axis, float min_dist, int ignore_deg,
                       double* diag, double* rhs, int64_t* row_start, int32_t* col, double* val);

/* Right-preconditioned BiCGSTAB (van der Vorst 1992) with the Jacobi
 * preconditioner on the system above, from the initial guess x (in/out),
 * until ||r|| <= tol ||b|| or max_iter iterations; returns ||r|| / ||b||
 * (the solver error QuadraticPlacer records, quadratic_placer.cpp:130-134).
 * *iters receives the iterations run.  Dot products are summed in the fixed
 * order of the GPU's block reduction, so iterates compare bit for bit. */
double ora_bicgstab(int n, const double* diag, const int64_t* row_start, const int32_t* col,
                    const double* val, const double* rhs, double* x, double tol, int max_iter,
                    int* iters);

typedef struct ora_qp_config {
    int32_t max_iter;          /* qp_max_iter (outer) */
    float tol;                 /* qp_tol (BiCGSTAB relative residual) */
    int32_t solver_max_iter;   /* qp_solver_max_iter */
    float outer_tol;           /* qp_outer_tol (relative HPWL change) */
    float min_distance;        /* qp_min_distance */
    int32_t ignore_net_degree;
} ora_qp_config;

/* QuadraticPlacer::place (quadratic_placer.cpp:149-228): movable nodes start
 * at the core centre; each outer iteration takes the bound pins of the
 * current positions, solves x then y (warm start = previous solution),
 * clamps every node inside the core and records HPWL and max(err_x, err_y);
 * it stops after iteration 5 once HPWL changes by less than outer_tol
 * relative, or after max_iter iterations.  x, y hold all nodes (fixed ones
 * are read); rows (may be NULL) get hpwl and err per iteration, at most
 * max_rows.  Returns the iterations run. */
int ora_qp_place(const ora_design* d, const float* net_weight, const ora_qp_config* cfg,
                 float* x, float* y, double* hpwl_rows, double* err_rows, int max_rows);

/* OpenMP threads the oracle uses (0 = leave runtime default). */
void ora_set_threads(int n);
int ora_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
