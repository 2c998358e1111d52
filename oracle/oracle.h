/*
 * oracle.h -- CPU restatement of BharatOpt's first-order (PDHG) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing here is part of the shipped library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so, and only as the checker (or the timed CPU baseline),
 * never as the thing measured on the GPU or as a fallback.
 *
 * Contents
 *   orc_csr_from_coo      src/sparse_matrix.cpp:59-110   (COO -> CSR: sort, dup-sum, zero-prune)
 *   orc_csr_spmv          src/sparse_matrix.cpp:153-163  (row-order dot products)
 *   orc_csr_spmv_t        src/sparse_matrix.cpp:175-185  (row-order scatter A^T y)
 *   orc_pdhg_ref          src/first_order_solver.cpp:103-279 (plain diagonal PDHG, FP64,
 *                          the reference's loop order; built with -ffp-contract=off)
 *   orc_kkt               rel-KKT of min c'x, A_E x=b_E, A_I x<=b_I, l<=x<=u (FP64);
 *                          the reference has no KKT measure -- definition follows
 *                          PDLP (Applegate et al. 2021), pinned by closed forms and by
 *                          HiGHS optimal primal/dual pairs in tests/
 *   orc_max_violation     src/first_order_solver.cpp:202-212 (max row / bound violation)
 *   orc_pdlp_omp          the restarted Halpern PDHG of the GPU engine in FP64 on
 *                          all host cores (no reference counterpart; the multi-core
 *                          CPU baseline of the new algorithm; libpdlp_omp.so)
 *
 * Parity pins (tests/test_oracle_*.py): the reference's own KATs
 * (tests/test_first_order_solver.cpp, tests/test_sparse_matrix.cpp,
 * tests/test_gpu_foundation.cpp), the 24,700-iteration Benchmark-10 figure
 * (docs/phase-12-report.md:92), and bit-for-bit agreement with oracle/_ref (the
 * reference's own sources compiled by oracle/Makefile) on seeded instances.
 */
#ifndef BO_ORACLE_H
#define BO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sparse ------------------------------------------------------------ */
/* COO -> CSR.  Entries are ordered by (row, col); duplicates are summed in
 * input order; exact zeros (after summing) are dropped.  Output arrays must
 * hold rows+1 / nnz_in entries.  Returns the resulting nnz, or -1 on an out
 * of range index (the reference throws std::out_of_range). */
long orc_csr_from_coo(int rows, int cols, long nnz_in, const int* ri, const int* ci,
                      const double* v, int* row_ptr, int* col_idx, double* val);

void orc_csr_spmv(int rows, const int* rp, const int* ci, const double* v,
                  const double* x, double* y);
void orc_csr_spmv_t(int rows, int cols, const int* rp, const int* ci, const double* v,
                    const double* x, double* y);

/* ---- reference PDHG ------------------------------------------------------ */
typedef struct {
    long   max_iterations;      /* FirstOrderSolverOptions::max_iterations (50000) */
    double feasibility_tol;     /* 1e-7 */
    double primal_step;         /* 0.01, used when adaptive == 0 */
    double dual_step;           /* 0.01 */
    int    adaptive;            /* 1: tau_j = 0.95/colsum, sigma_i = 0.95/rowsum */
    long   check_frequency;     /* capped at 100 like the reference */
} orc_pdhg_opts;

typedef struct {
    int    status;              /* 0 OPTIMAL, 1 MAX_ITERATIONS, 2 NUMERICAL_FAILURE */
    long   iterations;          /* reference leaves 0 on NUMERICAL_FAILURE; we report the
                                   loop counter in loop_iterations */
    long   loop_iterations;
    double constraint_violation;
    double bound_violation;
    int    used_average;        /* 1 when the ergodic average was returned */
} orc_pdhg_out;

/* Standard form (reference convention): min c'x s.t. row i: a_i x <= b_i, or
 * a_i x == b_i when is_eq[i]; lb <= x <= ub (+-inf allowed).  x_out[n],
 * y_out[m].  Returns status. */
int orc_pdhg_ref(int m, int n, const int* rp, const int* ci, const double* v,
                 const double* b, const double* c, const double* lb, const double* ub,
                 const unsigned char* is_eq, const orc_pdhg_opts* o,
                 double* x_out, double* y_out, orc_pdhg_out* out);

/* ---- FP64 verifier ------------------------------------------------------- */
typedef struct {
    double r_primal;        /* || [A_E x - b_E ; (A_I x - b_I)^+] ||_2              */
    double r_dual;          /* || lambda - proj_Lambda(lambda) ||_2 (+ ||min(y_I,0)||) */
    double primal_obj;      /* c'x                                                    */
    double dual_obj;        /* -b'y + sum_j l_j plam_j^+ - u_j plam_j^-               */
    double gap_rel;         /* |pobj-dobj| / (1 + |pobj| + |dobj|)                    */
    double b_norm, c_norm;
    double rel_primal;      /* r_primal / (1 + ||b||)                                 */
    double rel_dual;        /* r_dual   / (1 + ||c||)                                 */
    double rel_kkt;         /* max(rel_primal, rel_dual, gap_rel)                     */
    double max_row_viol;    /* inf-norm row violation                                 */
    double max_bound_viol;  /* inf-norm bound violation                               */
} orc_kkt_out;

void orc_kkt(int m, int n, const int* rp, const int* ci, const double* v,
             const double* b, const double* c, const double* lb, const double* ub,
             const unsigned char* is_eq, const double* x, const double* y,
             orc_kkt_out* out);

/* ---- the new algorithm on all host cores (pdlp_omp.c, libpdlp_omp.so) ---- */
typedef struct {
    long long max_iterations;
    double eps_rel_kkt;     /* stop when the FP64 rel-KKT (orc_kkt's definition) <= this */
    int check_every;        /* steps between checks (64) */
    int ruiz_iterations;    /* 10 */
    int pock_chambolle;     /* 1 */
    int threads;            /* OpenMP threads; <= 0: all */
    double eps_objective;   /* > 0: also stop only when |c'x - dual obj| <= this (1 + |c'x|) */
    const double* x_start;  /* optional, n: starting primal point (unscaled; projected on the bounds) */
    const double* y_start;  /* optional, m: starting duals (unscaled; inequality duals clipped at 0) */
    double primal_weight_start; /* > 0: starting primal weight instead of ||c_hat|| / ||b_hat|| */
    const char* checkpoint_path;  /* optional: the iteration state is written there every
                                     checkpoint_every iterations (at a check), and a run that finds a
                                     matching file there continues from it instead of starting */
    long long checkpoint_every;
} orc_pdlp_opts;

typedef struct {
    int status;             /* 0 rel-KKT reached, 1 iteration limit, 2 numerical failure */
    long long iterations;
    int restarts;
    double rel_kkt, rel_primal, rel_dual, rel_gap, primal_obj, dual_obj;
    int threads;
    double step_size;       /* eta = 0.95 / ||A_hat||_2 */
    double primal_weight;   /* omega at the end */
    double prep_s;          /* CSC copy of the matrix (a one-time layout cost) */
    double setup_s;         /* prep + scaling + power iteration */
    double solve_s;         /* the whole call, setup included */
} orc_pdlp_out;

/* Restarted, reflected Halpern PDHG in FP64 with OpenMP; x_out[n], y_out[m]
 * (unscaled, the better of z and T(z) at the last check).  Returns status. */
int orc_pdlp_omp(int m, int n, const int* rp, const int* ci, const double* v, const double* b,
                 const double* c, const double* lb, const double* ub, const unsigned char* is_eq,
                 const orc_pdlp_opts* o, double* x_out, double* y_out, orc_pdlp_out* out);

/* Reference-style violation measure (src/first_order_solver.cpp:202-212). */
void orc_max_violation(int m, int n, const int* rp, const int* ci, const double* v,
                       const double* b, const double* lb, const double* ub,
                       const unsigned char* is_eq, const double* x,
                       double* max_row_viol, double* max_bound_viol);

#ifdef __cplusplus
}
#endif
#endif
