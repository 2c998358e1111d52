/*
 * bo_capi.h -- C ABI of the MI355X (gfx950) PDHG engine, libbo_hip.so.
 *
 * Plain C: pointers, sizes and POD structs, no C++ types, no exceptions.
 * Every entry point names the reference interface it replaces (paths are
 * relative to the BharatOpt source tree).  bo_status ordinals equal
 * bharatopt::GPUStatus (include/bharatopt/gpu_backend.hpp:15-22).
 *
 * Memory: "d_" pointers are device pointers obtained from bo_buf_alloc;
 * every other pointer is host memory owned by the caller.  Handles
 * (bo_pdlp*) own their device memory until destroyed.  Calls on one handle
 * must not overlap; distinct handles may be used from distinct host threads.
 */
#ifndef BO_CAPI_H
#define BO_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bo_status {
    BO_SUCCESS = 0,
    BO_GPU_UNAVAILABLE = 1,
    BO_GPU_UNINITIALIZED = 2,
    BO_OUT_OF_MEMORY = 3,
    BO_INVALID_ARGUMENT = 4,
    BO_EXECUTION_FAILED = 5
} bo_status;

/* GPUDeviceInfo (include/bharatopt/gpu_backend.hpp:24-32) plus CU count. */
typedef struct bo_device_info {
    int is_available;
    char device_name[256];
    size_t total_memory_bytes;
    size_t free_memory_bytes;
    int compute_capability_major; /* gfx9.5.0 -> 9 / 50 */
    int compute_capability_minor;
    int multiprocessor_count;
    char status_message[256];
} bo_device_info;

/* GpuTimingResult (include/bharatopt/gpu_backend.hpp:34-41); kernel time from HIP events. */
typedef struct bo_timing {
    double h2d_time_ms;
    double kernel_time_ms;
    double d2h_time_ms;
    double sync_time_ms;
    double total_gpu_time_ms;
} bo_timing;

/* Last error text of the calling thread ("" if none). */
const char* bo_last_error(void);

/* ---- device and buffers --------------------------------------------------- */
/* GpuBackend::initialize (gpu_backend.hpp:175, src/gpu_backend.cpp:14-32). */
bo_status bo_device_init(bo_device_info* info);
/* GpuBuffer<T>::allocate (gpu_backend.hpp:78-93).  bytes==0 -> *d_ptr=NULL, SUCCESS. */
bo_status bo_buf_alloc(size_t bytes, void** d_ptr);
/* GpuBuffer<T>::deallocate (gpu_backend.hpp:95-99). */
void bo_buf_free(void* d_ptr);
/* GpuBuffer<T>::copy_host_to_device / copy_device_to_host (gpu_backend.hpp:101-120). */
bo_status bo_buf_h2d(void* d_dst, const void* src, size_t bytes);
bo_status bo_buf_d2h(void* dst, const void* d_src, size_t bytes);

/* ---- FP64 primitives of GpuBackend (gpu_backend.hpp:179-227) -------------- */
/* GpuBackend::spmv: y = A x, CSR, each row summed in column order (bit-equal
 * to CSRMatrix::multiply, src/sparse_matrix.cpp:153-163). */
bo_status bo_spmv_csr_f64(size_t rows, size_t cols, size_t nnz, const int* d_row_ptr,
                          const int* d_col_idx, const double* d_val, const double* d_x,
                          double* d_y, bo_timing* timing);
/* GpuBackend::vector_add / vector_sub / vector_scale / axpy (src/gpu_backend.cpp:135-269). */
bo_status bo_vec_add_f64(const double* d_a, const double* d_b, double* d_out, size_t n, bo_timing* t);
bo_status bo_vec_sub_f64(const double* d_a, const double* d_b, double* d_out, size_t n, bo_timing* t);
bo_status bo_vec_scale_f64(const double* d_in, double alpha, double* d_out, size_t n, bo_timing* t);
bo_status bo_axpy_f64(double alpha, const double* d_x, const double* d_y_in, double* d_y_out,
                      size_t n, bo_timing* t);
/* GpuBackend::dot_product / norm2 (src/gpu_backend.cpp:271-327); fixed-order tree. */
bo_status bo_dot_f64(const double* d_a, const double* d_b, double* result, size_t n, bo_timing* t);
bo_status bo_norm2_f64(const double* d_v, double* result, size_t n, bo_timing* t);

/* Measurement utility (no reference counterpart): `reps` launches of a plain
 * 16-byte-vector copy kernel over `bytes` bytes (read once, written once) on
 * the default stream; *ms is the mean launch time from HIP events.  Used to
 * calibrate hardware byte counters against a kernel of known traffic. */
bo_status bo_copy_kernel_calibrate(size_t bytes, int reps, double* ms);

/* ---- standard-form LP ----------------------------------------------------- */
/* min c'x s.t. row i: a_i x <= b_i (a_i x == b_i when is_eq[i]), lb <= x <= ub.
 * This is the form CPUFirstOrderSolver::solve builds (src/first_order_solver.cpp:68-101).
 * CSR with strictly ascending columns per row; all arrays are host memory. */
typedef struct bo_lp {
    int m;
    int n;
    long long nnz;
    const int* row_ptr;
    const int* col_idx;
    const double* val;
    const double* b;
    const double* c;
    const double* lb;
    const double* ub;
    const unsigned char* is_eq;
} bo_lp;

/* ---- reference-semantics PDHG (FP64) -------------------------------------- */
/* Options of FirstOrderSolverOptions used by the loop (first_order_solver.hpp:29-38). */
typedef struct bo_ref_options {
    long long max_iterations;
    double feasibility_tolerance;
    double primal_step_size;
    double dual_step_size;
    int adaptive_step_size;
    long long check_frequency;
} bo_ref_options;

typedef struct bo_ref_result {
    int status; /* FirstOrderSolverStatus: 0 OPTIMAL, 1 MAX_ITERATIONS, 2 NUMERICAL_FAILURE */
    long long iterations;
    long long loop_iterations;
    double constraint_violation;
    double bound_violation;
    int used_average;
    double kernel_time_ms;
    double transfer_time_ms;
    int loop; /* how the loop ran: 0 kernel per pass, 1 one workgroup (matrix in memory),
                 2 one workgroup with the matrix resident in LDS */
} bo_ref_result;

/* The PDHG loop of CPUFirstOrderSolver::solve (src/first_order_solver.cpp:103-279)
 * on the GPU in FP64 with the reference's summation orders and no fused
 * multiply-adds: results are bit-equal to the reference.  x_out[n], y_out[m]. */
bo_status bo_pdhg_ref_solve(const bo_lp* lp, const bo_ref_options* opts, double* x_out,
                            double* y_out, bo_ref_result* result);
/* No counterpart in the reference.  For callers that cannot allocate memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): bo_pdhg_ref_solve with the same arguments and bit-identical
 * results, calling no hipMalloc / hipFree or other allocating routine.  Every
 * device array of the solve comes from one fixed array per GPU (about 2 MiB,
 * held until the process ends) and the host-side CSC from fixed static host
 * arrays.  Limit: m + n <= 7680 and nnz <= 65536 (the LPs that run in one
 * workgroup); a larger LP returns BO_INVALID_ARGUMENT with bo_last_error()
 * saying so.  Calls are serialized by a process-wide lock: on one GPU they
 * share that array, which keeps the last call's contents (each call clears
 * what must start at zero), so overlapping calls would otherwise race. */
bo_status bo_pdhg_ref_solve_noalloc(const bo_lp* lp, const bo_ref_options* opts, double* x_out,
                                    double* y_out, bo_ref_result* result);

/* ---- PDLP engine (FP32 iterates, FP64 checks) ----------------------------- */
typedef struct bo_pdlp_params {
    long long max_iterations;  /* default 100000 */
    double eps_rel_kkt;        /* stop when FP64 rel-KKT <= this (default 1e-6) */
    int ruiz_iterations;       /* default 10 */
    int pock_chambolle;        /* 1: l1 (alpha=1) rescaling after Ruiz */
    int check_every;           /* iterations between FP64 KKT / restart checks; 0 (default): 64 */
    int lanes_rows;            /* CSR pass schedule: k > 0 segment kernel with k nonzeros per lane,
                                  -k (1..64) tiled gather (LDS-resident vector block, SELL-64
                                  slices) + reduce kernels, cut for k tiles per CU,
                                  -c stream kernel with an LDS budget of c (1024/2048/4096)
                                  nonzeros per workgroup, 0 = autotune */
    int lanes_cols;            /* CSC pass schedule, same encoding; the batched engine uses the tiled
                                  layout for a pass whose code is <= 0 and the segment kernel
                                  otherwise */
    int fp64_iterates;         /* 0 (default): FP32 iterates and scaled matrix values; 1: FP64 both (the
                                  error comparison run); 2: FP64 iterates over FP32 matrix values, the
                                  multi-kernel loop only (single-LP engine; the batched and family
                                  engines take 2 as 1) */
    int use_graph;             /* 1: replay each check interval as a hipGraph */
    double time_limit_s;       /* 0 = none; else wall-clock limit of one solve call, tail rounds and FP64
                                  polish included: LPs still running when it passes stop with status 3.
                                  A time limit turns the one-workgroup solve off (single_workgroup) */
    int verbose;               /* diagnostics on stderr: >= 1 the single-LP engine's set-up phases and
                                  their wall times at create, >= 2 also the batched engine's progress
                                  per check interval and per tail round; 0 (default) silent */
    int hot_columns;           /* entries of x+ kept in LDS by the CSR pass: -1 tune, 0 off */
    long long stall_iterations; /* batched and family engines, FP32 iterates: an LP whose FP64 rel-KKT
                                   has not halved within max(this, iterations / 2) iterations stops
                                   as STALLED (4); 0 = never (default 16384: on the 512-LP ragged
                                   batch 4096 stalled 146 LPs that FP32 would have finished, at 8%
                                   more batch time) */
    int fp64_polish;           /* batched and family engines: finish every STALLED LP with FP64 iterates from
                                  its FP32 point, within its remaining iteration budget (default 1) */
    int compact_tail;          /* batched engine: once at most 1/4 of a batch's LPs still run, continue
                                  them in a compact batch (segment kernels), from their points, primal
                                  weights and remaining budgets, repeatedly (default 1) */
    int single_workgroup;      /* single-LP engine: 1 run the whole solve in one workgroup (no kernel
                                  launch per pass), 0 never, -1 (default) the tuner times both on LPs
                                  of up to 2^18 nonzeros without a time limit */
    double eps_objective;      /* > 0: an LP also stops only when |c'x - dual objective| <= this
                                  x (1 + |c'x|); with x and y feasible to eps_rel_kkt the optimum lies
                                  between the two objectives, so c'x is then within this (relative) of
                                  it.  0 (default): rel-KKT alone, whose gap term is relative to
                                  1 + |c'x| + |dual objective| */
} bo_pdlp_params;

typedef struct bo_pdlp_stats {
    int status; /* 0 OPTIMAL (rel-KKT met), 1 MAX_ITERATIONS, 2 NUMERICAL_FAILURE, 3 TIME_LIMIT,
                   4 STALLED (FP32 iterates stopped improving; see fp64_polish) */
    long long iterations;
    int restarts;
    int used_average;
    double rel_kkt;
    double rel_primal;
    double rel_dual;
    double rel_gap;
    double primal_obj;
    double dual_obj;
    double primal_weight;
    double step_size;
    double setup_ms;
    double solve_ms;
    double spmv_csr_bytes; /* algorithmic bytes of one A x product    */
    double spmv_csc_bytes; /* algorithmic bytes of one A' y product   */
    int lanes_rows;        /* schedule used by the CSR pass (encoding of bo_pdlp_params) */
    int lanes_cols;        /* ... and by the CSC pass */
    int hot_columns;       /* LDS-cached entries of x+ in the CSR pass (0 = none) */
    long long fp64_polish_iterations; /* of `iterations`, those run by the FP64 polish */
    int check_every;       /* check cadence used (single-LP engine) */
    int single_workgroup;  /* the solve ran in one workgroup (single-LP engine): 1 over the matrix in
                              memory, 2 with the scaled matrix resident in LDS */
} bo_pdlp_stats;

/* Default parameters. */
void bo_pdlp_default_params(bo_pdlp_params* p);
/* Upload + Ruiz/Pock-Chambolle scaling + step size.  Replaces the set-up part of
 * CPUFirstOrderSolver::solve (src/first_order_solver.cpp:101-156). */
bo_status bo_pdlp_create(const bo_lp* lp, const bo_pdlp_params* params, struct bo_pdlp** out);
/* Optional starting point in the original (unscaled) space; NULL = reference
 * initialisation (box midpoint / finite bound, y = 0). */
bo_status bo_pdlp_set_warm_start(struct bo_pdlp* h, const double* x0, const double* y0);
/* Iterate to rel-KKT <= eps (src/first_order_solver.cpp:158-247 replaced by PDLP). */
bo_status bo_pdlp_solve(struct bo_pdlp* h, bo_pdlp_stats* stats);
/* Solution in the original space (x[n], y[m], reference sign convention y_I >= 0). */
bo_status bo_pdlp_get_solution(struct bo_pdlp* h, double* x, double* y);
void bo_pdlp_destroy(struct bo_pdlp* h);
/* Time `reps` products of each SpMV kernel of the handle (HIP events, on the
 * handle's stream).  ms are per-launch averages. */
bo_status bo_pdlp_time_spmv(struct bo_pdlp* h, int reps, double* csr_ms, double* csc_ms);

/* Diagnostics without a counterpart in the reference: the schedule a pass
 * runs with and the algorithmic bytes of its kernels (pass 0: CSR pass,
 * A x+ and the dual step; pass 1: CSC pass, A'y and the primal step). */
typedef struct bo_pdlp_pass_info {
    int schedule;        /* encoding of bo_pdlp_params.lanes_rows */
    int kernels;         /* launches per pass: 1 fused, 2 tiled (gather, reduce) */
    int blocks;          /* tiled: blocks of the gathered vector */
    int block_entries;   /* tiled: entries per block, held in LDS */
    int tiles;           /* tiled: workgroups of the gather kernel */
    int slices;          /* tiled: SELL-64 slices */
    long long pairs;     /* tiled: (block, segment) pairs holding nonzeros */
    long long padded;    /* tiled: stored entries including slice padding */
    double gather_bytes; /* algorithmic bytes of one launch of the (gather) kernel */
    double reduce_bytes; /* ... of the reduce + epilogue kernel; 0 when fused */
    int reduce_loads;    /* tiled: partials loaded per round of the reduction (tuned) */
} bo_pdlp_pass_info;
bo_status bo_pdlp_get_pass_info(struct bo_pdlp* h, int pass, bo_pdlp_pass_info* out);
/* Per-launch averages over `reps` alternating iterations, HIP events on the
 * handle's stream: ms4 = [CSC gather, CSC reduce, CSR gather, CSR reduce]
 * (a fused pass reports its kernel as gather and 0 as reduce). */
bo_status bo_pdlp_time_kernels(struct bo_pdlp* h, int reps, double* ms4);

/* ---- batched PDLP: many LPs of uneven size solved together ---------------- */
/* No single counterpart in the reference, which solves one model per
 * FirstOrderLPSolver::solve call (src/first_order_solver.cpp:311-319): this is
 * the pre-built-CSR batch entry point for callers with many independent LPs.
 * The LPs are laid out block-diagonally on the device and every launch covers
 * all of them; each LP keeps its own step sizes, primal weight, restarts and
 * FP64 stopping test (bo_pdlp_params applies to every LP) and stops on its
 * own.  Per-LP results equal a bo_pdlp solve of that LP up to summation order. */
typedef struct bo_pdlp_batch_stats {
    int count;                 /* LPs in the batch */
    int solved;                /* LPs that met eps_rel_kkt */
    long long max_iterations;  /* iterations of the slowest LP */
    long long sum_iterations;  /* sum over LPs */
    long long loop_steps;      /* steps the batch loop ran */
    double solve_ms;           /* host wall time of bo_pdlp_batch_solve */
    int tiled;                 /* 1: both passes on the tiled schedule */
    int polished;              /* LPs finished by the FP64 polish */
    int tail_lps;              /* LPs handed to the tail rounds (compact batches, FP64 polish) */
    double tail_ms;            /* host wall time of the tail rounds (set-up included), part of solve_ms */
} bo_pdlp_batch_stats;
struct bo_pdlp_batch;
bo_status bo_pdlp_batch_create(const bo_lp* lps, int count, const bo_pdlp_params* params,
                               struct bo_pdlp_batch** out);
/* Warm start for every LP: x and y concatenated in batch order (sum n, sum m
 * entries, original space); NULL, NULL clears it. */
bo_status bo_pdlp_batch_set_warm_start(struct bo_pdlp_batch* h, const double* x, const double* y);
/* Starting primal weight of every LP (count entries, > 0; 0 keeps the default
 * ||c_hat|| / ||b_hat|| of the scaled LP); NULL clears them.  Part of a warm
 * start: the weight a solve of a similar LP ended with (bo_pdlp_stats.primal_weight). */
bo_status bo_pdlp_batch_set_primal_weights(struct bo_pdlp_batch* h, const double* w);
/* New costs c (sum n entries) and / or right-hand sides b (sum m entries) for
 * every LP, batch order, original space; NULL keeps that part.  The matrices,
 * bounds, row senses and the device layout stay, so a parametric family of
 * LPs over one matrix is re-solved without building the batch again. */
bo_status bo_pdlp_batch_set_objective_rhs(struct bo_pdlp_batch* h, const double* c, const double* b);
/* per_lp (count entries, may be NULL): the stats of every LP, as a single solve reports
 * them; their solve_ms is the batch's wall time. */
bo_status bo_pdlp_batch_solve(struct bo_pdlp_batch* h, bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total);
/* Solution of LP k in its original space (x[n_k], y[m_k]). */
bo_status bo_pdlp_batch_get_solution(struct bo_pdlp_batch* h, int k, double* x, double* y);
/* Mean time (ms, HIP events on the batch's stream) of each iteration kernel with every LP running,
 * on the state the last solve left: ms4 = [CSC gather, CSC reduce + primal step, CSR gather,
 * CSR reduce + dual step]; bytes4 = their algorithmic bytes.  Needs the tiled layout of both passes. */
bo_status bo_pdlp_batch_time_kernels(struct bo_pdlp_batch* h, int reps, double* ms4, double* bytes4);
void bo_pdlp_batch_destroy(struct bo_pdlp_batch* h);

/* ---- parametric family: one matrix, up to 64 LPs ------------------------- */
/* No counterpart in the reference, which solves one LP per call
 * (src/first_order_solver.cpp:15-291).  LPs that share A, bounds and row
 * senses and differ in c and b are solved together: vectors are stored
 * LP-minor, one wavefront per row / column with one LP per lane, so each
 * matrix entry is read once per pass for all of them.  The scaling and the
 * step-size estimate depend on A alone and are computed once, at create.
 * Arrays are host memory; c and x are count x n, b and y count x m, row-major
 * (one row per LP).  Statistics and the FP64 polish are those of the batched
 * engine. */
struct bo_pdlp_family;
/* The family of `lp`'s matrix; room for `capacity` (1..64) LPs, all starting as lp. */
bo_status bo_pdlp_family_create(const bo_lp* lp, int capacity, const bo_pdlp_params* params,
                                struct bo_pdlp_family** out);
/* The next solves' LPs: count (1..capacity) cost and right-hand-side rows;
 * both NULL keeps the current ones (count unchanged).  Clears the warm start. */
bo_status bo_pdlp_family_set_lps(struct bo_pdlp_family* h, int count, const double* c, const double* b);
/* Starting points (count x n, count x m, original space); NULL, NULL clears them. */
bo_status bo_pdlp_family_set_warm_start(struct bo_pdlp_family* h, const double* x, const double* y);
/* Starting primal weights (count values, finite, >= 0; 0 for an LP means its
 * own ||c_hat|| / ||b_hat||), e.g. the weights solves of similar LPs ended
 * with (bo_pdlp_stats.primal_weight); NULL clears them.  Kept until changed. */
bo_status bo_pdlp_family_set_primal_weights(struct bo_pdlp_family* h, const double* w);
bo_status bo_pdlp_family_solve(struct bo_pdlp_family* h, bo_pdlp_stats* per_lp, bo_pdlp_batch_stats* total);
/* Solutions of the last solve, original space (count x n, count x m; either may be NULL). */
bo_status bo_pdlp_family_get_solutions(struct bo_pdlp_family* h, double* x, double* y);
/* Mean time (ms, HIP events on the engine's stream) of one primal product
 * pass (CSC, x update) and one dual pass (CSR, y update) over all `count`
 * LPs, run on the state the last solve left; ms2[0] primal, ms2[1] dual. */
bo_status bo_pdlp_family_time_passes(struct bo_pdlp_family* h, int reps, double* ms2);
void bo_pdlp_family_destroy(struct bo_pdlp_family* h);

/* ---- warm-start network (BF16 MLP on the matrix cores) -------------------- */
/* No counterpart in the reference, which has no learned component.  A
 * multilayer perceptron n_in -> width -> ... -> n_out (`layers` products,
 * ReLU between them) with BF16 weights and activations, FP32 accumulation,
 * FP32 master weights and Adam; it maps LP parameter changes to primal/dual
 * warm starts for bo_pdlp_set_warm_start / bo_pdlp_batch_set_warm_start.
 * All arrays are host memory, row-major, `batch` rows. */
struct bo_mlp;
bo_status bo_mlp_create(int n_in, int width, int n_out, int layers, int max_batch, unsigned long long seed,
                        struct bo_mlp** out);
/* Number of parameters: for each layer l, W_l (d_in x d_out, row-major) then b_l (d_out). */
long long bo_mlp_num_params(const struct bo_mlp* h);
bo_status bo_mlp_get_params(struct bo_mlp* h, float* params);
bo_status bo_mlp_set_params(struct bo_mlp* h, const float* params); /* also resets Adam */
bo_status bo_mlp_forward(struct bo_mlp* h, const float* x, int batch, float* y);
/* Mean squared error of forward(x) against target, and its gradient (layout of the parameters). */
bo_status bo_mlp_gradients(struct bo_mlp* h, const float* x, const float* target, int batch, double* loss,
                           float* grads);
/* One Adam step (beta 0.9 / 0.999) at learning rate lr; *loss before the step. */
bo_status bo_mlp_train_step(struct bo_mlp* h, const float* x, const float* target, int batch, float lr,
                            double* loss);
/* Mean time (ms, HIP events on the network's stream) of one forward pass
 * (what 0) or forward + backward pass (what 1) over `batch` rows of the
 * device-resident input (the last one uploaded, zeros at first). */
bo_status bo_mlp_time(struct bo_mlp* h, int what, int batch, int reps, double* ms);
void bo_mlp_destroy(struct bo_mlp* h);
/* The product kernel alone: C[M x N] = op(A) op(B), BF16 operands given as raw
 * 16-bit words, A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], FP32 C. */
bo_status bo_gemm_bf16(int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                       const unsigned short* B, long long sbk, long long sbn, float* C);

/* ---- full drop-in solve on a flattened LPModel ----------------------------- */
/* LPModel (include/bharatopt/lp_model.hpp:60-127) as arrays: row-wise terms in
 * insertion order. con_sense: 0 LE, 1 EQ, 2 GE, 3 RANGED; var_type: 0 continuous,
 * 1 integer, 2 binary (NULL = continuous); sense: 0 minimize, 1 maximize. */
typedef struct bo_model {
    int num_vars;
    int num_cons;
    int sense;
    double obj_offset;
    const double* lb;
    const double* ub;
    const double* obj;
    const int* var_type;
    const int* con_sense;
    const double* rhs;
    const double* range_upper;
    const long long* term_ptr;
    const int* term_var;
    const double* term_val;
} bo_model;

/* FirstOrderSolverOptions (first_order_solver.hpp:29-38) + engine choice. */
typedef struct bo_fo_options {
    long long max_iterations;
    double feasibility_tolerance;
    double optimality_tolerance;
    double primal_step_size;
    double dual_step_size;
    int adaptive_step_size;
    long long check_frequency;
    int engine; /* 0: reference PDHG semantics (FP64, bit-equal); 1: PDLP to rel-KKT */
    double eps_rel_kkt;
    int cpu_backend; /* 1: CPUFirstOrderSolver, the reference's loop on the host (engine ignored);
                        0: GPUFirstOrderSolver */
    double eps_objective; /* PDLP engine: bo_pdlp_params.eps_objective (0: rel-KKT alone) */
} bo_fo_options;

/* FirstOrderSolverResult / Stats (first_order_solver.hpp:41-60). */
typedef struct bo_fo_result {
    int status;
    long long iterations;
    double objective_value;
    double primal_objective;
    double dual_objective;
    double primal_residual;
    double dual_residual;
    double constraint_violation;
    double bound_violation;
    double solve_time_ms;
    double transfer_time_ms;
    double kernel_time_ms;
    long long primal_len;
    long long dual_len;
    char message[256];
    /* SolutionVerifier::verify of the returned primal solution against the
     * original model with the reported objective, at the options'
     * feasibility tolerance (1e-7 by default); status -1 when there is no
     * solution to verify, else bo_verify_result.status */
    int verify_status;
    double verify_max_violation; /* largest constraint or bound violation the verifier found */
} bo_fo_result;

/* PresolveEngine::presolve (include/bharatopt/presolve.hpp:96-102), host only:
 * status (PresolveStatus ordinal), reduced sizes, objective offset and the
 * reduced model's nonzero count.  Any pointer may be NULL. */
bo_status bo_presolve_stats(const bo_model* model, int* status, int* reduced_vars,
                            int* reduced_cons, double* obj_offset, long long* reduced_nnz);

/* FirstOrderLPSolver::solve (first_order_solver.hpp:103-113, src/first_order_solver.cpp:311-319):
 * validate -> presolve -> standard form -> PDHG -> postsolve -> objective, the
 * loop on the GPU (GPUFirstOrderSolver::solve, :295-306) or, with
 * cpu_backend = 1, on the host (CPUFirstOrderSolver::solve, :14-290).
 * primal_out holds num_vars, dual_out num_cons entries.  Returns
 * BO_GPU_UNAVAILABLE when the GPU backend was asked for and no GPU is visible. */
bo_status bo_first_order_solve(const bo_model* model, const bo_fo_options* opts,
                               bo_fo_result* result, double* primal_out, double* dual_out);

/* ---- routing cost model ---------------------------------------------------- */
/* Predicted wall time of one GPU first-order solve on the MI355X. */
typedef struct bo_cost_estimate {
    double total_ms;          /* create_ms + fixed_ms + iterations x per_iteration_ms */
    double create_ms;         /* upload and layout (PDLP: handle creation with schedule tuning) */
    double fixed_ms;          /* per-solve device set-up */
    double per_iteration_ms;
    long long iterations;     /* iterations assumed */
    int loop;                 /* reference engine: its loop tier (bo_ref_result.loop); PDLP: -1 */
    int in_domain;            /* 1: nnz within the measured range (1,500 .. 17.1M) */
} bo_cost_estimate;

/* Replaces the GPU estimate of AnalyticalCostEstimator::estimate
 * (src/cost_estimator.cpp:188-197, the analytic 24.5 + 0.0008 nnz ms) with
 * costs measured on an MI355X (calibrate_cost.py), interpolated in log(nnz).
 * engine 0: reference PDHG (bo_pdhg_ref_solve), 1: PDLP (bo_pdlp_*);
 * iterations 0: 50,000 (reference) or 10,000 (PDLP).  Host arithmetic only. */
bo_status bo_pdhg_cost_estimate(int m, int n, long long nnz, int engine, long long iterations,
                                bo_cost_estimate* out);

/* ---- solution verifier ----------------------------------------------------- */
/* VerificationOptions (include/bharatopt/solution_verifier.hpp:13-18). */
typedef struct bo_verify_options {
    double feasibility_tolerance; /* default 1e-7 */
    double integrality_tolerance; /* default 1e-5 */
    double objective_tolerance;   /* default 1e-4 */
    int check_objective;          /* default 1 */
} bo_verify_options;

/* VerificationResult (solution_verifier.hpp:20-40) without its texts.
 * status: 0 verified, 1 failed verification, 2 objective mismatch (the
 * reported objective differs), 3 dimension mismatch; diagnostics: the number
 * of diagnostic messages the verifier produced. */
typedef struct bo_verify_result {
    int verified;
    int has_reported_objective;
    double objective_reported;
    double objective_recomputed;
    double objective_difference;
    double maximum_constraint_violation;
    double maximum_lower_bound_violation;
    double maximum_upper_bound_violation;
    double maximum_integrality_violation;
    long long violated_constraint_count;
    long long violated_bound_count;
    long long violated_integrality_count;
    long long variable_count;
    long long constraint_count;
    int status;
    int diagnostics;
} bo_verify_result;

/* SolutionVerifier::verify (solution_verifier.hpp:46-52, src/solution_verifier.cpp:42-215),
 * host FP64: bounds, integrality, row activities (LE / GE / EQ / RANGED) and
 * the objective of candidate x (x_len entries; a length other than
 * model->num_vars is the verifier's dimension failure) against the original
 * model.  reported_objective != NULL: the overload that also compares a
 * reported objective.  opts NULL: the defaults.  Needs no GPU. */
bo_status bo_verify_solution(const bo_model* model, const double* x, long long x_len,
                             const double* reported_objective, const bo_verify_options* opts,
                             bo_verify_result* result);

/* ---- MPS reader ------------------------------------------------------------ */
/* MpsParser::parse_string / parse_file (include/bharatopt/mps_parser.hpp:40-49,
 * src/mps_parser.cpp:44-529) into a model handle whose flattened LPModel
 * (bo_model above) feeds bo_first_order_solve. */
typedef struct bo_mps_info {
    int status; /* MpsParseStatus ordinal: 0 SUCCESS, 1 FILE_NOT_FOUND, 2 MISSING_ENDATA,
                   3 UNKNOWN_SECTION, 4 MALFORMED_ROW, 5 MALFORMED_COLUMN, 6 MALFORMED_BOUND,
                   7 UNCLOSED_INTEGER_MARKER, 8 INVALID_NUMERIC_VALUE */
    int num_vars;
    int num_cons;
    long long nonzeros; /* coefficients placed (MpsParseResult::nonzeros_parsed) */
    int integer_vars;
    int binary_vars;
    double parse_time_ms;
    char problem_name[128];
    char message[256];
} bo_mps_info;

struct bo_mps;
/* text != NULL: parse that text; else read the file at `path`.  Returns
 * BO_SUCCESS whenever the reader ran: info->status tells whether the text was
 * a model, and *model is set (non-NULL) only when it was. */
bo_status bo_mps_read(const char* text, const char* path, bo_mps_info* info, struct bo_mps** model);
/* The model as a bo_model whose arrays belong to the handle (valid until
 * bo_mps_destroy). */
bo_status bo_mps_model(const struct bo_mps* model, bo_model* view);
void bo_mps_destroy(struct bo_mps* model);

#ifdef __cplusplus
}
#endif
#endif
