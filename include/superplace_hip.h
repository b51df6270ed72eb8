/* superPlace-HIP: C ABI of the MI355X (gfx950) backend for the ePlace
 * mGP/cGP iteration.
 *
 * Plain pointers and sizes only; no HIP, torch or C++ types cross the
 * boundary.  Every function returns an sp_hip_status; on failure
 * sp_hip_last_error() gives a thread-local message.  Each declaration names
 * the reference interface it replaces (file:line in the reference tree).
 *
 * Threading: an sp_hip_poisson or sp_hip_gp handle is used by one thread at a
 * time (the reference calls PoissonBackend::solve from one thread per
 * instance, poisson_backend.h:29-37); separate handles may be used
 * concurrently, each owns its own HIP stream.
 */
#ifndef SUPERPLACE_HIP_H
#define SUPERPLACE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sp_hip_status {
    SP_HIP_OK = 0,
    SP_HIP_ENODEV = 1,      /* no usable GPU: the caller falls back to its CPU backend,
                               as makePoissonBackend does when the CUDA factory returns
                               nullptr (src/numeric/poisson_factory.cpp:14-21) */
    SP_HIP_EINVAL = 2,      /* bad argument (std::runtime_error in the reference,
                               src/numeric/cpu/cpu_poisson.cpp:78-83, :220-223) */
    SP_HIP_EOOM = 3,        /* device allocation failed */
    SP_HIP_EDEVICE = 4,     /* HIP runtime error */
    SP_HIP_ENONFINITE = 5,  /* non-finite coordinate / gradient / metric
                               (InvariantError "non_finite", src/gp/eplace.cpp:140-142, :192-197) */
    SP_HIP_EINTERNAL = 6
} sp_hip_status;

enum {
    SP_HIP_FLAG_FP64 = 1,    /* FP64 arithmetic everywhere (the plain FP64 path) */
    SP_HIP_FLAG_BF16X3 = 2,  /* Poisson handles and sp_hip_gemm: the dense transforms as
                                split-BF16 products on the BF16 matrix cores (each FP32
                                operand as hi + lo bfloat16, three products); FP32 only */
    SP_HIP_FLAG_DENSE = 4    /* Poisson handles: the dense transforms on every grid size
                                (power-of-two grids otherwise take the FFT passes) */
};

/* Thread-local message of the last failing call on this thread. */
const char* sp_hip_last_error(void);
/* Number of visible HIP devices (0 without a GPU). */
int sp_hip_device_count(int* count);

/* ---------------------------------------------------------------- Poisson
 * Replaces PoissonBackend (src/numeric/poisson_backend.h:20-46) and its
 * factory makePoissonBackend(BackendConfig) (poisson_backend.h:48-60,
 * src/numeric/poisson_factory.cpp:14-21).  Solves lap(phi) = -rho with
 * Neumann boundaries and returns the field xi = -grad(phi) (not a gradient),
 * with the conventions of src/numeric/cpu/cpu_poisson.cpp:3-40.  Grids are
 * row-major with y fastest, data[x*ny + y] (src/numeric/bin_grid_span.h:6-19).
 * All buffers are allocated at create time; solve does not allocate. */
typedef struct sp_hip_poisson sp_hip_poisson;

/* BackendConfig{nx, ny, step_x, step_y}; flags: SP_HIP_FLAG_FP64, or
 * SP_HIP_FLAG_DENSE and / or SP_HIP_FLAG_BF16X3 (FP32) to choose the transforms.
 * nx, ny >= 2 (any size: the transforms are padded internally).
 * Returns SP_HIP_ENODEV when no GPU is present. */
int sp_hip_poisson_create(int nx, int ny, float step_x, float step_y, int flags,
                          sp_hip_poisson** out);
/* PoissonBackend::solve(density, phi, field_x, field_y) on host spans of
 * nx*ny floats (cpu_poisson.cpp:149-207); synchronous. */
int sp_hip_poisson_solve(sp_hip_poisson* p, const float* density, float* phi, float* field_x,
                         float* field_y);
/* Device-resident variant: density_dev holds nx*ny values and out3_dev
 * 3*nx*ny values (phi | field_x | field_y), float (or double with
 * SP_HIP_FLAG_FP64); enqueued on the handle's stream, then synchronised.
 * No counterpart in the reference (its CUDA backend always crosses to the
 * host, cuda_poisson.cu:194-210); it is the data path of the engine below. */
int sp_hip_poisson_solve_dev(sp_hip_poisson* p, const void* density_dev, void* out3_dev);
/* sp_hip_poisson_solve_dev ordered on the caller's stream (a hipStream_t;
 * NULL = the default stream) instead of synchronised: the solve runs after
 * the work enqueued on `stream` so far, and work enqueued on `stream`
 * afterwards runs after the solve; the host does not wait.  The device-span
 * form of PoissonBackend::solve (src/numeric/poisson_backend.h:36-37) for a
 * caller whose density and fields live on its own stream, where the
 * reference's CUDA backend enqueues on a stream of its own and synchronises
 * (src/numeric/cuda/cuda_poisson.cu:194-210).  One call at a time per
 * handle, as every handle call. */
int sp_hip_poisson_solve_async(sp_hip_poisson* p, const void* density_dev, void* out3_dev,
                               void* stream);
/* The FP32 solve of sp_hip_poisson_solve_dev without a handle and without
 * allocating: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  nx, ny: powers of two from 64 to 1024
 * (SP_HIP_EINVAL otherwise); density_dev nx*ny floats, out3_dev 3*nx*ny
 * floats (phi | field_x | field_y, data[x*ny + y]); enqueued on `stream`
 * (a hipStream_t; NULL = the default stream), not synchronised.  The
 * intermediates live in a static device array of the library (12.6 MB per
 * GPU, held until the process ends) that every call on one GPU shares:
 * calls that overlap on one GPU race.  Results equal those of
 * sp_hip_poisson_solve_dev (FP32 handle) on the same grid. */
int sp_hip_poisson_solve_noalloc(int nx, int ny, float step_x, float step_y,
                                 const float* density_dev, float* out3_dev, void* stream);
/* PoissonBackend::name(): "hip-fft" (power-of-two grids: LDS FFT passes) or
 * "hip-mfma" (other sizes: dense transforms on the matrix cores), with
 * "-fp64" appended on the FP64 path. */
const char* sp_hip_poisson_name(const sp_hip_poisson* p);
void sp_hip_poisson_destroy(sp_hip_poisson* p);

/* ---------------------------------------------------------------- GP engine
 * Device-resident evaluation and iteration of one placement stage: the data
 * of PlaceDB (src/db/place_db.h:27-71) and the loop of EPlace::run
 * (src/gp/eplace.cpp:223-526) with NesterovOptimizer
 * (src/numeric/cpu/nesterov.cpp:51-240). */

typedef struct sp_hip_gp sp_hip_gp;

/* Pure SoA + CSR database, node order [movable | fixed | filler], centre
 * coordinates.  node2pin_start has num_nodes + num_fillers + 1 entries
 * (fillers own no pins, src/gp/filler.cpp:129-130).  Rows give the core
 * (their bounding box) and the base density (bin_grid.cpp:252-281). */
typedef struct sp_hip_design {
    int32_t num_movable, num_nodes, num_fillers, num_pins, num_nets, num_rows;
    const float* node_x;
    const float* node_y;
    const float* node_w;
    const float* node_h;
    const uint8_t* node_flags;   /* 1 macro, 2 fixed, 4 filler, 8 terminal_NI */
    const float* pin_offset_x;
    const float* pin_offset_y;
    const int32_t* pin2node;
    const int32_t* pin2net;
    const int32_t* net2pin_start;
    const int32_t* flat_net2pin;
    const int32_t* node2pin_start;
    const int32_t* flat_node2pin;
    const float* row_ly;
    const float* row_height;
    const float* row_lx;
    const float* row_hx;
    float core_lx, core_ly, core_hx, core_hy;
} sp_hip_design;

/* The GP fields of Config (src/util/config.h:77-118, configs/default.json). */
typedef struct sp_hip_gp_config {
    float target_density;
    float target_overflow;
    float cgp_target_overflow;
    int32_t gp_max_iter;
    int32_t ignore_net_degree;
    int32_t gp_bin_dim;            /* 0 = automatic power of two (bin_grid.cpp:190-204) */
    float delta_hpwl_ref;
    int32_t lambda_relative;       /* lambda_schedule == "relative" */
    double delta_hpwl_relative;
    int32_t use_bb;
    int32_t gp_backtrack;
    int32_t filler_only_iters;
    int32_t gp_stagnation_window;
    float gp_stagnation_min_gain;
    int32_t stage;                 /* 0 mGP, 1 FILLERONLY, 2 cGP (src/gp/eplace.h:24-28) */
    int32_t precision;             /* 32: FP32 bulk + FP64 reductions; 64: FP64 everywhere */
    int32_t max_backtrack;         /* OptConfig::max_backtrack (src/numeric/optimizer.h:223) */
    /* How the iteration loop is driven (no counterpart in the reference):
     * 0 = host decisions, the host reads the scalars after every evaluation;
     * 1 = decisions on the device, kernels launched slot by slot;
     * 2 = decisions on the device, graph_slots iterations per hipGraph launch
     * (0 = the launch shapes' graph_slots, below). */
    int32_t loop_mode;
    int32_t graph_slots;
} sp_hip_gp_config;

/* Launch shapes of the device loop per bin-grid size: the choices the
 * autotuner (superplace_amd/tune.py) makes from measured timings.  No shape
 * changes a result bit: the scatter sums integers, the sort only reorders
 * the density records, the column pass transforms every column pair alike
 * and the graph only groups iterations.  The library holds a compiled-in
 * table (the tuner's last output, superplace_amd/shapes.json); calls below
 * override it for the instances created afterwards.  No counterpart in the
 * reference (fixed CPU code paths). */
typedef struct sp_hip_shapes {
    int32_t scatter_chunk;   /* density records per scatter block: 256, 512, 1024 or 2048 */
    int32_t sort_period;     /* slots between the sort passes' launches, 1-8 */
    int32_t col_pairs;       /* column pairs per block of the Poisson column pass: 1, 2, 4
                              * or 8; 0 = the column pass shares its launch with the
                              * WA blocks (power-of-two grids) */
    int32_t graph_slots;     /* iterations per graph launch where the configuration's
                              * graph_slots is 0, 1-64 */
} sp_hip_shapes;
/* bin_dim 0 sets the shapes of every size without an entry of its own. */
int sp_hip_shapes_set(int bin_dim, const sp_hip_shapes* shapes);
/* The shapes an instance with this many bins per side takes. */
int sp_hip_shapes_get(int bin_dim, sp_hip_shapes* shapes);
/* Back to the compiled-in table. */
int sp_hip_shapes_reset(void);

/* One metrics.csv row (src/util/metrics.h:12-28). */
typedef struct sp_hip_iter_metrics {
    int32_t iter;
    int32_t evaluations;
    double hpwl, overflow, lambda, gamma, step_size, grad_norm_wl, grad_norm_den;
    double eval_lambda, eval_gamma, delta_hpwl, lambda_mu, elapsed_ms;
} sp_hip_iter_metrics;

/* GpResult (src/gp/eplace.h:32-42) plus the stage's initial values. */
typedef struct sp_hip_gp_result {
    int32_t iterations, converged, stagnated, rollback_iter;
    int64_t evaluations;
    double hpwl, overflow;
    double initial_hpwl, initial_overflow, initial_lambda;
    double seconds_loop;
} sp_hip_gp_result;

/* FP64 scalars of one objAndGrad evaluation (eplace.cpp:136-203). */
typedef struct sp_hip_eval_scalars {
    double tau, wl, hpwl, sum_wl, sum_den, dx2, dg2, sy, gg, nonfinite;
} sp_hip_eval_scalars;

/* Optional host outputs of one evaluation (any pointer may be NULL):
 * bin grids dim*dim (node_q / filler_q: the raw fixed-point sums, value =
 * q * 2^-frac_bits); force/grad_wl/grad 2*n_opt in pos order (x half, y half). */
typedef struct sp_hip_eval_debug {
    int64_t* node_q;
    int64_t* filler_q;
    double* node_density;
    double* filler_density;
    double* rho;
    double* field_x;
    double* field_y;
    double* force;
    double* grad_wl;
    double* grad;
} sp_hip_eval_debug;

typedef struct sp_hip_gp_info {
    int32_t dim;              /* bins per side */
    int32_t n_opt;            /* optimised nodes (pos has 2*n_opt entries) */
    int32_t frac_bits;        /* fixed-point fraction bits of the density scatter */
    int32_t precision;
    double step_x, step_y, origin_x, origin_y;
    double movable_area_scaled;   /* tau denominator (bin_grid.cpp:182-188) */
    int64_t device_bytes;         /* device memory the instance allocated so far */
    sp_hip_shapes shapes;         /* the launch shapes it runs with */
} sp_hip_gp_info;

/* Kernel families timed with HIP events on the engine's stream in the
 * device-driven loop (enabled by sp_hip_gp_set_profiling, which makes
 * loop_mode 2 launch its slots directly, as loop_mode 1 does): every kernel
 * of a family carries its own start/stop event pair (hipExtLaunchKernel, the
 * dispatch's own timestamps), and one launch of the family counts the sum of
 * its kernels' durations.  Each family is one kernel, except POISSON (the
 * transform products of one solve) and SORT (the two passes of the spatial
 * sort). */
enum { SP_HIP_K_SCATTER = 0, SP_HIP_K_FINALIZE, SP_HIP_K_POISSON, SP_HIP_K_WA_NETS,
       SP_HIP_K_NODE_GRAD, SP_HIP_K_UPDATE, SP_HIP_K_SORT, SP_HIP_K_REDUCE, SP_HIP_K_COUNT };

/* Builds the device copy of the design and the bin grid for one stage
 * (EPlace::EPlace + BinGrid::initialize, eplace.cpp:205-221).  The copy is
 * kept in an order of the engine's own (fillers by position; cells by
 * connectivity when the design's small nets join far-apart indices); every
 * entry point below takes and returns nodes in the caller's numbering. */
int sp_hip_gp_create(const sp_hip_design* d, const sp_hip_gp_config* cfg, sp_hip_gp** out);
int sp_hip_gp_info_get(const sp_hip_gp* gp, sp_hip_gp_info* info);
/* The optimised node indices (caller numbering) in the engine's pos order,
 * the order of the per-node vectors of sp_hip_eval_debug: buildOptSet's set
 * (eplace.cpp:82-94), its order the engine's. */
int sp_hip_gp_opt_index(const sp_hip_gp* gp, int32_t* idx);
/* One objAndGrad at host positions x, y of ALL nodes (written as given,
 * like EPlace's scatter), with lambda and gamma.  dbg may be NULL. */
int sp_hip_gp_eval(sp_hip_gp* gp, const float* x, const float* y, float lambda, float gamma,
                   sp_hip_eval_scalars* out, sp_hip_eval_debug* dbg);
/* EPlace::run(stage): runs the stage to convergence, stagnation or the
 * iteration limit.  rows may be NULL; at most max_rows rows are written. */
int sp_hip_gp_run(sp_hip_gp* gp, sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int max_rows);
/* Many placement stages run together (no counterpart in the reference,
 * which places one design per process): count FP32 instances (power-of-two
 * bin grids: the automatic rule or a power-of-two gp_bin_dim; any sizes, and
 * grids of different sizes in one batch) advance through their stages in
 * batched launches, each kernel of an iteration launched once for all of
 * them (the Poisson passes once per grid size), `slots` iterations per
 * hipGraph launch (<= 0: 8) until every stage has ended.  results[i] and
 * the positions (x[i], y[i]: all nodes of design i; x, y or their entries
 * may be NULL) equal those of sp_hip_gp_run on the same design and
 * configuration, bit for bit. */
int sp_hip_gp_run_batch(int count, const sp_hip_design* designs, const sp_hip_gp_config* cfgs,
                        int slots, sp_hip_gp_result* results, float* const* x, float* const* y);
/* Current node positions (all nodes; after run: the delivered placement). */
int sp_hip_gp_get_positions(const sp_hip_gp* gp, float* x, float* y);
/* mask: bit f (SP_HIP_K_*) times kernel family f with HIP events; 0 = off. */
int sp_hip_gp_set_profiling(sp_hip_gp* gp, int mask);
int sp_hip_gp_kernel_stats(const sp_hip_gp* gp, double* ms, int64_t* launches, double* bytes,
                           double* flops);
void sp_hip_gp_destroy(sp_hip_gp* gp);

/* ---------------------------------------------------------------- BookShelf
 * Replaces BookshelfReader (src/io/bookshelf_reader.h:20-41: read at :23,
 * readPlacement at :26) and BookshelfWriter::writePl
 * (src/io/bookshelf_writer.h:15).  Host code; no GPU is needed.  The reader
 * fills the same database as the reference: node order [movable | fixed],
 * centres, macros = movable nodes taller than 1.5 row heights, net->pin and
 * node->pin CSR.  Errors (missing file, unknown node in a net, malformed
 * number) return SP_HIP_EINVAL with the message. */
typedef struct sp_bookshelf sp_bookshelf;

int sp_bookshelf_read(const char* aux_path, sp_bookshelf** out);
/* A view of the database for sp_hip_gp_create (num_fillers = 0); the
 * pointers stay valid until sp_bookshelf_destroy. */
int sp_bookshelf_design(const sp_bookshelf* b, sp_hip_design* view);
/* Per node orientation code (N E S W FN FE FS FW = 0..7), per net weight,
 * per row site step and site count, and the row height (most frequent). */
int sp_bookshelf_extra(const sp_bookshelf* b, const uint8_t** orient,
                       const float** net_weight, const float** row_step,
                       const int32_t** row_sites, float* row_height);
/* which = 0: node names, 1: net names, each followed by '\n', into buf of
 * cap bytes; *need = bytes required (buf may be NULL to ask). */
int sp_bookshelf_names(const sp_bookshelf* b, int which, char* buf, int64_t cap, int64_t* need);
/* readPlacement: positions (and /FIXED flags) from another .pl, by name. */
int sp_bookshelf_read_pl(sp_bookshelf* b, const char* pl_path);
/* writePl: the num_nodes node centres x, y as lower-left corners, with the
 * reference's exact format (bookshelf_writer.cpp:16-37). */
int sp_bookshelf_write_pl(const sp_bookshelf* b, const char* path, const float* x,
                          const float* y);
/* writePl of a database whose flags and orientations changed since reading
 * (mLG fixes macros and mirrors flipped ones): num_nodes flags (bit 2 writes
 * /FIXED) and BookShelf orientation codes; either may be NULL (as read). */
int sp_bookshelf_write_pl_state(const sp_bookshelf* b, const char* path, const float* x,
                                const float* y, const uint8_t* node_flags,
                                const uint8_t* node_orient);
void sp_bookshelf_destroy(sp_bookshelf* b);

/* ---------------------------------------------------- Quadratic placement
 * Replaces QuadraticPlacer::place (src/init/quadratic_placer.h:18,
 * src/init/quadratic_placer.cpp:149-228), the Kraftwerk2 Bound2Bound initial
 * placement: movable nodes start at the core centre; each outer iteration
 * takes the bound pins of the current positions (computeNetBoundPins,
 * src/db/hpwl.cpp:70-102), builds both axes' B2B systems (pairs of
 * forEachB2BPair, src/init/b2b_model.h:56-82), solves them with
 * Jacobi-preconditioned BiCGSTAB from the previous solution, clamps every
 * node into the core and records HPWL; it stops after iteration 5 once HPWL
 * moves by less than outer_tol relative, or after max_iter iterations.  On
 * the GPU, FP64 solver.  Needs a GPU (SP_HIP_ENODEV without one). */
typedef struct sp_qp_config {     /* the qp_* fields of Config, src/util/config.h:49-74 */
    int32_t max_iter;             /* qp_max_iter (20) */
    float tol;                    /* qp_tol (1e-6): BiCGSTAB relative residual */
    int32_t solver_max_iter;      /* qp_solver_max_iter (100) */
    float outer_tol;              /* qp_outer_tol (1e-4) */
    float min_distance;           /* qp_min_distance (1.0) */
    int32_t ignore_net_degree;    /* ignore_net_degree (100): larger nets are left out */
} sp_qp_config;
/* Positions in: the design's node_x / node_y (fixed nodes); out: x, y with
 * num_nodes entries.  net_weight (num_nets) may be NULL (all 1).  rows (may
 * be NULL) receive iter, hpwl, grad_norm_wl = max(err_x, err_y) and
 * elapsed_ms per outer iteration, at most max_rows. */
int sp_qp_place(const sp_hip_design* d, const float* net_weight, const sp_qp_config* cfg,
                float* x, float* y, sp_hip_iter_metrics* rows, int max_rows,
                int32_t* iterations);
/* The B2B system of one axis (0 = x, 1 = y) at the design's positions, as
 * the GPU assembles it (assemble, quadratic_placer.cpp:43-113): per movable
 * node i a row of entries (movable node, -w) in row_start[i] ..
 * row_start[i+1], FP64 diagonal and right-hand side.  Call with diag NULL to
 * get *nnz only; cap is the capacity of col / val. */
int sp_qp_b2b_system(const sp_hip_design* d, const float* net_weight, int axis,
                     float min_distance, int32_t ignore_net_degree, double* diag, double* rhs,
                     int64_t* row_start, int32_t* col, double* val, int64_t cap, int64_t* nnz);

/* ---------------------------------------------------------------- Serving
 * Many placement jobs on one GPU.  No counterpart in the reference (it
 * places one design per process): a scheduler for callers that place many
 * designs.  Each job is one sp_hip_gp-style instance (one stage: the
 * config's stage, from the design's positions) with its design and
 * optimizer state resident on the device; running jobs advance together,
 * each on its own stream with its device-driven loop (loop_mode 1 or 2)
 * kept two launches deep, so their kernels overlap.  A job is admitted, in
 * arrival order, once its arrival time has passed, fewer than max_active
 * jobs run and its footprint fits in what the budget has left; the host
 * set-up of the next job overlaps the running ones.  Every job's result
 * equals that of sp_hip_gp_run on the same design and configuration, bit
 * for bit.  One host thread drives a server; times are seconds on the
 * server's clock (0 at create). */
enum { SP_HIP_JOB_QUEUED = 0, SP_HIP_JOB_RUNNING = 1, SP_HIP_JOB_DONE = 2, SP_HIP_JOB_FAILED = 3 };
typedef struct sp_hip_job_info {
    int32_t state;          /* SP_HIP_JOB_* */
    int32_t status;         /* sp_hip_status of a failed job (SP_HIP_EOOM: footprint above the budget) */
    int64_t device_bytes;   /* footprint: the estimate until set up, then the measured bytes */
    double arrival_s, admit_s, finish_s;   /* -1 until reached */
} sp_hip_job_info;
typedef struct sp_hip_server sp_hip_server;
/* Device bytes an instance of this design and configuration allocates (an
 * upper bound of what it measures once set up). */
int sp_hip_gp_footprint(const sp_hip_design* d, const sp_hip_gp_config* cfg, int64_t* bytes);
int sp_hip_server_create(int64_t hbm_budget, int32_t max_active, sp_hip_server** out);
/* Copies the design; the job becomes eligible at arrival_s. */
int sp_hip_server_submit(sp_hip_server* s, const sp_hip_design* d, const sp_hip_gp_config* cfg,
                         double arrival_s, int32_t* job);
/* One scheduling round (never waits for the device); any pointer may be NULL. */
int sp_hip_server_pump(sp_hip_server* s, double* now_s, int32_t* active, int32_t* queued,
                       int32_t* finished);
/* Pumps until every submitted job is done or failed (SP_HIP_EINVAL after timeout_s). */
int sp_hip_server_drain(sp_hip_server* s, double timeout_s);
int sp_hip_server_now(const sp_hip_server* s, double* now_s);
/* A job's state, result and (once done) the positions of all its nodes;
 * res, x, y may be NULL. */
int sp_hip_server_job(const sp_hip_server* s, int32_t job, sp_hip_job_info* info,
                      sp_hip_gp_result* res, float* x, float* y);
/* Frees a finished job's host copy of its design and positions. */
int sp_hip_server_release(sp_hip_server* s, int32_t job);
void sp_hip_server_destroy(sp_hip_server* s);

/* -------------------------------------------------------------- Legalizer
 * Replaces legalizeCells and buildAbacusSegments (src/lg/abacus.h:34-35,
 * src/lg/abacus.cpp:108-286): Abacus standard-cell legalization over the
 * site rows, bit for bit the reference's result.  Host code; no GPU is
 * needed (the algorithm is one sequential greedy pass).  row_step holds the
 * site width of every row (sp_bookshelf_extra gives it); the rows' hx is
 * lx + step * sites.  x, y: the centres of all num_nodes nodes, in and out
 * (only standard cells move).  Invalid rows or non-finite geometry return
 * SP_HIP_EINVAL (std::invalid_argument there). */
enum { SP_ABACUS_FULL_TRIALS = 1,   /* AbacusOptions::referenceTrials */
       SP_ABACUS_AREA_WEIGHT = 2 }; /* AbacusOptions::areaWeight */
typedef struct sp_abacus_result {   /* AbacusResult, src/lg/abacus.h:21-32 */
    int32_t precondition_ok, cells, failed, rows, segments, rounded_widths;
    int32_t selfcheck_overlaps, selfcheck_offsite;
    double disp_mean, disp_max;
} sp_abacus_result;
int sp_abacus_legalize(const sp_hip_design* d, const float* row_step, float* x, float* y,
                       int flags, sp_abacus_result* result);
/* The free row segments at the design's positions; call with seg_row NULL
 * to get *count only. */
int sp_abacus_segments(const sp_hip_design* d, const float* row_step, int32_t* seg_row,
                       float* seg_lx, float* seg_hx, int32_t cap, int32_t* count);

/* Replaces legalizeMacros (src/lg/macro_sa.h:40, src/lg/macro_sa.cpp:493-1272):
 * simulated annealing over the movable macros (flag macro, not fixed, below
 * num_movable) followed by greedy repairs, bit for bit the reference's
 * result on the same design and seed.  Host code; no GPU is needed (one
 * sequential Markov chain).  site_width is the first row's site step,
 * row_height the design's row height (<= 0 means 1).  In/out: x, y (centres
 * of all num_nodes nodes; only macros move), node_flags (macros gain flag
 * 2, fixed), pin_offset_x (negated on the pins of horizontally flipped
 * macros), orient (BookShelf orientation codes, N<->FN and S<->FS on flip;
 * may be NULL).  net_weight may be NULL (all 1).  sink (may be NULL)
 * receives the summaries mlg_<step>_{overlap, disp_mean, disp_max,
 * worst_macro, original_x, original_y, x, y} for the steps sa, separate1,
 * chain1, local_grid, separate2, chain2, unlimited_grid (when run),
 * postlegal and final, mlg_postlegal_accepted, and the mLG duration;
 * worst_macro is node_names[node] (the node index when node_names is
 * NULL).  A spatial index that disagrees with a full scan (checked up to
 * 8000 macros) returns SP_HIP_EINVAL (InvariantError there). */
typedef struct sp_mlg_config {   /* the macro_sa_* fields of Config, src/util/config.h */
    uint64_t random_seed;        /* Config::random_seed (1002) */
    int32_t moves;               /* macro_sa_moves; 0 = max(20000, 300 x overlapping) */
    float init_accept;           /* macro_sa_init_accept (0.9) */
    float temp_ratio;            /* macro_sa_temp_ratio (1e-4) */
    float disp_weight;           /* macro_sa_disp_weight (1.0) */
} sp_mlg_config;
typedef struct sp_mlg_result {   /* MacroSaResult, src/lg/macro_sa.h:20-35 */
    int32_t num_macros, moves, accepted, repaired, chained, hot_macros;
    double overlap_before, overlap_after_sa, overlap_after, sa_ms;
    double t0, typical_worsen, hpwl_before, hpwl_after;
} sp_mlg_result;
int sp_macro_legalize(const sp_hip_design* d, const float* net_weight, float site_width,
                      float row_height, const sp_mlg_config* cfg, float* x, float* y,
                      uint8_t* node_flags, float* pin_offset_x, uint8_t* orient,
                      struct sp_metrics* sink, const char* const* node_names,
                      sp_mlg_result* result);

/* ---------------------------------------------------------------- Metrics
 * Replaces MetricsSink (src/util/metrics.h:40-90, src/util/metrics.cpp:19-140):
 * per-iteration rows to <out_dir>/metrics.csv (rewritten whenever a stage
 * ends) and a summary.json of string keys, number keys and durations_ms,
 * byte for byte in the reference's formats.  Host code; no GPU is needed.
 * Stages nest (mGP inside gp): rows go to the innermost open stage.
 * Non-finite rows or numbers return SP_HIP_ENONFINITE (InvariantError
 * "non_finite" in the reference); a file that cannot be written returns
 * SP_HIP_EINVAL (the reference only warns, so a caller may ignore it). */
typedef struct sp_metrics sp_metrics;
/* out_dir may be NULL or "" (no metrics.csv); created if missing. */
int sp_metrics_create(const char* out_dir, sp_metrics** out);
int sp_metrics_begin_stage(sp_metrics* m, const char* stage);
/* MetricsSink::push for n rows; max_displacement (n values) may be NULL (0). */
int sp_metrics_push(sp_metrics* m, const sp_hip_iter_metrics* rows, int n,
                    const double* max_displacement);
/* MetricsSink::endStage: records the stage's duration, rewrites metrics.csv. */
int sp_metrics_end_stage(sp_metrics* m, double elapsed_ms);
int sp_metrics_record_duration(sp_metrics* m, const char* label, double ms);
int sp_metrics_set_number(sp_metrics* m, const char* key, double value);
int sp_metrics_set_string(sp_metrics* m, const char* key, const char* value);
/* The summary keys EPlace::run writes for a stage (src/gp/eplace.cpp:266-270,
 * :509-513): <stage>_initial_hpwl, _initial_overflow, _initial_lambda,
 * _hpwl_reference, _restarts (when restarts >= 0), _evaluations, _converged,
 * _stagnated, _rollback_iter. */
int sp_metrics_gp_result(sp_metrics* m, const char* stage, const sp_hip_gp_result* r,
                         double hpwl_reference, int restarts);
int sp_metrics_write_summary(const sp_metrics* m, const char* path);
void sp_metrics_destroy(sp_metrics* m);

/* ------------------------------------------------------------------ Fillers
 * Replaces initializeFillers (src/gp/filler.h:29, filler.cpp:47-139) and
 * scatterFillers (filler.h:32, filler.cpp:33-45) for a caller that holds the
 * design as arrays: the library cannot grow the caller's arrays, so the
 * plan (FillerInfo, filler.h:18-25) and the positions are two calls, and
 * the caller appends `count` nodes of width x height with flag 4 (filler)
 * after its fixed nodes.  Host code; no GPU is needed.  Results are
 * bit-identical to the reference's on the same design and seed. */
typedef struct sp_filler_info {
    int32_t count;
    float width, height;
    double whitespace_area, node_area_scaled, total_filler_area;
} sp_filler_info;
/* The plan for a design without fillers (d->num_fillers is ignored; only
 * nodes below num_nodes count) at target density (> 0). */
int sp_filler_plan(const sp_hip_design* d, float target_density, sp_filler_info* info);
/* Centres of `count` fillers of width x height drawn uniformly inside the
 * design's core with the reference's generator (std::mt19937 seeded with
 * `seed`, i.e. Config::random_seed + the caller's offset; a float uniform
 * draw for x, then for y, per filler). */
int sp_filler_scatter(const sp_hip_design* d, int count, float width, float height,
                      uint64_t seed, float* x, float* y);

/* `count` successive draws of the reference's Rng::uniform(lo, hi)
 * (src/util/rng.h:15-18: std::mt19937 seeded with `seed`, one
 * uniform_real_distribution<float>(lo, hi) draw each), the generator of the
 * reference's test fixtures (test/test_eplace.cpp:45-121).  Host code. */
int sp_rng_uniform(uint64_t seed, int64_t count, float lo, float hi, float* out);

/* Batched C = A * B^T of row-major matrices (A m x k, B n x k, C m x n; dims
 * multiples of 64) on the matrix cores, the primitive of the Poisson
 * transforms.  Device pointers, float (or double with SP_HIP_FLAG_FP64);
 * SP_HIP_FLAG_BF16X3: float, as split-BF16 products.  Synchronous.  No
 * counterpart in the reference (FFTW / cuFFT there). */
int sp_hip_gemm(int flags, int batch, const void* const* a, const void* const* b,
                void* const* c, int m, int n, int k);

#ifdef __cplusplus
}
#endif
#endif
