/*
 * pdhg_ref.c -- FP64 restatement of the reference PDHG and its sparse kernels.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off so
 * that every product and sum rounds exactly as the reference's x86-64 build.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* std::max / std::min / bharatopt::clamp with their exact NaN behaviour
 * (include/bharatopt/config.hpp:25-28 for clamp). */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

/* ---- COO -> CSR (src/sparse_matrix.cpp:59-110) --------------------------- */
long orc_csr_from_coo(int rows, int cols, long nnz_in, const int* ri, const int* ci,
                      const double* v, int* row_ptr, int* col_idx, double* val) {
    for (long k = 0; k < nnz_in; ++k) {
        if (ri[k] < 0 || (rows > 0 && ri[k] >= rows) || ci[k] < 0 || (cols > 0 && ci[k] >= cols))
            return -1;
    }
    /* bucket by row (stable), then order each row by column (stable insertion
     * sort keeps duplicates in input order). */
    long* cnt = (long*)calloc((size_t)rows + 1, sizeof(long));
    long* perm = (long*)malloc((size_t)(nnz_in > 0 ? nnz_in : 1) * sizeof(long));
    for (long k = 0; k < nnz_in; ++k) cnt[ri[k] + 1]++;
    for (int i = 0; i < rows; ++i) cnt[i + 1] += cnt[i];
    long* pos = (long*)malloc(((size_t)rows + 1) * sizeof(long));
    memcpy(pos, cnt, ((size_t)rows + 1) * sizeof(long));
    for (long k = 0; k < nnz_in; ++k) perm[pos[ri[k]]++] = k;
    for (int i = 0; i < rows; ++i) {
        for (long a = cnt[i] + 1; a < cnt[i + 1]; ++a) {
            long key = perm[a];
            long b = a - 1;
            while (b >= cnt[i] && ci[perm[b]] > ci[key]) { perm[b + 1] = perm[b]; --b; }
            perm[b + 1] = key;
        }
    }
    long out = 0;
    row_ptr[0] = 0;
    for (int i = 0; i < rows; ++i) {
        long a = cnt[i];
        while (a < cnt[i + 1]) {
            int col = ci[perm[a]];
            double s = v[perm[a]];
            long b = a + 1;
            while (b < cnt[i + 1] && ci[perm[b]] == col) { s += v[perm[b]]; ++b; }
            if (s != 0.0) { col_idx[out] = col; val[out] = s; ++out; }
            a = b;
        }
        row_ptr[i + 1] = (int)out;
    }
    free(cnt); free(perm); free(pos);
    return out;
}

/* ---- SpMV (src/sparse_matrix.cpp:153-163, 175-185) ----------------------- */
void orc_csr_spmv(int rows, const int* rp, const int* ci, const double* v,
                  const double* x, double* y) {
    for (int i = 0; i < rows; ++i) {
        double sum = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) sum += v[k] * x[ci[k]];
        y[i] = sum;
    }
}

void orc_csr_spmv_t(int rows, int cols, const int* rp, const int* ci, const double* v,
                    const double* x, double* y) {
    for (int j = 0; j < cols; ++j) y[j] = 0.0;
    for (int i = 0; i < rows; ++i) {
        double xi = x[i];
        for (int k = rp[i]; k < rp[i + 1]; ++k) y[ci[k]] += v[k] * xi;
    }
}

/* max violation of rows (reference :202-206) */
static double row_violation(int m, const double* ax, const double* b, const unsigned char* is_eq) {
    double mv = 0.0;
    for (int i = 0; i < m; ++i) {
        double viol = (!is_eq[i]) ? smax(0.0, ax[i] - b[i]) : fabs(ax[i] - b[i]);
        mv = smax(mv, viol);
    }
    return mv;
}

void orc_max_violation(int m, int n, const int* rp, const int* ci, const double* v,
                       const double* b, const double* lb, const double* ub,
                       const unsigned char* is_eq, const double* x,
                       double* max_row_viol, double* max_bound_viol) {
    double* ax = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    orc_csr_spmv(m, rp, ci, v, x, ax);
    *max_row_viol = row_violation(m, ax, b, is_eq);
    double bv = 0.0;
    for (int j = 0; j < n; ++j) {
        if (x[j] < lb[j] - 1e-12) bv = smax(bv, lb[j] - x[j]);
        if (x[j] > ub[j] + 1e-12) bv = smax(bv, x[j] - ub[j]);
    }
    *max_bound_viol = bv;
    free(ax);
}

/* ---- PDHG main loop (src/first_order_solver.cpp:103-279) ------------------ */
int orc_pdhg_ref(int m, int n, const int* rp, const int* ci, const double* v,
                 const double* b, const double* c, const double* lb, const double* ub,
                 const unsigned char* is_eq, const orc_pdhg_opts* o,
                 double* x_out, double* y_out, orc_pdhg_out* out) {
    const size_t nn = (size_t)(n > 0 ? n : 1), mm = (size_t)(m > 0 ? m : 1);
    double* col_sums = (double*)calloc(nn, sizeof(double));
    double* row_sums = (double*)calloc(mm, sizeof(double));
    double* tau = (double*)malloc(nn * sizeof(double));
    double* sigma = (double*)malloc(mm * sizeof(double));
    double* x = (double*)malloc(nn * sizeof(double));
    double* y = (double*)calloc(mm, sizeof(double));
    double* x_hat = (double*)calloc(nn, sizeof(double));
    double* x_prev = (double*)malloc(nn * sizeof(double));
    double* at_y = (double*)calloc(nn, sizeof(double));
    double* a_xhat = (double*)calloc(mm, sizeof(double));
    double* x_sum = (double*)calloc(nn, sizeof(double));
    double* x_avg = (double*)malloc(nn * sizeof(double));
    double* ax = (double*)malloc(mm * sizeof(double));

    memset(out, 0, sizeof(*out));

    /* step sizes :104-128 -- sums accumulate in row order */
    for (int i = 0; i < m; ++i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            double a = fabs(v[k]);
            row_sums[i] += a;
            col_sums[ci[k]] += a;
        }
    }
    for (int j = 0; j < n; ++j) tau[j] = o->primal_step;
    for (int i = 0; i < m; ++i) sigma[i] = o->dual_step;
    if (o->adaptive) {
        for (int j = 0; j < n; ++j) tau[j] = (col_sums[j] > 1e-12) ? (0.95 / col_sums[j]) : 1.0;
        for (int i = 0; i < m; ++i) sigma[i] = (row_sums[i] > 1e-12) ? (0.95 / row_sums[i]) : 1.0;
    }

    /* initial point :131-156 */
    for (int j = 0; j < n; ++j) {
        double target = 0.0;
        if (lb[j] > -INFINITY && ub[j] < INFINITY) target = 0.5 * (lb[j] + ub[j]);
        else if (lb[j] > -INFINITY) target = lb[j];
        else if (ub[j] < INFINITY) target = ub[j];
        x[j] = sclamp(target, lb[j], ub[j]);
        x_prev[j] = x[j];
    }
    long avg_count = 0;
    long iter = 0;
    int converged = 0;
    long check_freq = o->check_frequency < 100 ? o->check_frequency : 100;
    if (check_freq < 1) check_freq = 1;  /* the reference divides by zero here */
    int status = 1;

    while (iter < o->max_iterations) {
        iter++;
        orc_csr_spmv_t(m, n, rp, ci, v, y, at_y);                  /* :163 */
        double max_primal_delta = 0.0;                             /* :166-175 */
        for (int j = 0; j < n; ++j) {
            double grad = c[j] + at_y[j];
            double next_x = x[j] - tau[j] * grad;
            double new_x = sclamp(next_x, lb[j], ub[j]);
            max_primal_delta = smax(max_primal_delta, fabs(new_x - x[j]));
            x[j] = new_x;
            x_sum[j] += x[j];
        }
        avg_count++;
        for (int j = 0; j < n; ++j) x_hat[j] = 2.0 * x[j] - x_prev[j];   /* :178-180 */
        orc_csr_spmv(m, rp, ci, v, x_hat, a_xhat);                      /* :183 */
        double max_dual_delta = 0.0;                                     /* :186-192 */
        for (int i = 0; i < m; ++i) {
            double next_y = y[i] + sigma[i] * (a_xhat[i] - b[i]);
            double new_y = (!is_eq[i]) ? smax(0.0, next_y) : next_y;
            max_dual_delta = smax(max_dual_delta, fabs(new_y - y[i]));
            y[i] = new_y;
        }
        memcpy(x_prev, x, (size_t)n * sizeof(double));                  /* :194 */

        if (iter % check_freq == 0 || iter == o->max_iterations) {     /* :197-246 */
            orc_csr_spmv(m, rp, ci, v, x, ax);
            double max_viol = row_violation(m, ax, b, is_eq);
            double max_bound_viol = 0.0;
            for (int j = 0; j < n; ++j) {
                if (x[j] < lb[j] - 1e-12) max_bound_viol = smax(max_bound_viol, lb[j] - x[j]);
                if (x[j] > ub[j] + 1e-12) max_bound_viol = smax(max_bound_viol, x[j] - ub[j]);
            }
            for (int j = 0; j < n; ++j) x_avg[j] = sclamp(x_sum[j] / (double)avg_count, lb[j], ub[j]);
            orc_csr_spmv(m, rp, ci, v, x_avg, ax);
            double avg_viol = row_violation(m, ax, b, is_eq);
            double best_viol = smin(max_viol, avg_viol);
            out->constraint_violation = best_viol;
            out->bound_violation = max_bound_viol;
            if (isnan(best_viol) || best_viol > 1e15) {
                status = 2;
                out->loop_iterations = iter;
                out->iterations = 0;
                goto done;
            }
            int feasible = (best_viol <= o->feasibility_tol && max_bound_viol <= o->feasibility_tol);
            int stationary = (max_primal_delta <= 1e-7 && max_dual_delta <= 1e-7);
            if (feasible && stationary && iter >= 100) { converged = 1; break; }
        }
    }
    out->iterations = iter;
    out->loop_iterations = iter;
    status = converged ? 0 : 1;

    /* iterate selection :262-279 */
    if (avg_count > 0) {
        for (int j = 0; j < n; ++j) x_avg[j] = sclamp(x_sum[j] / (double)avg_count, lb[j], ub[j]);
    } else {
        memcpy(x_avg, x, (size_t)n * sizeof(double));
    }
    orc_csr_spmv(m, rp, ci, v, x_avg, ax);
    {
        double avg_viol = row_violation(m, ax, b, is_eq);
        int use_avg = (!converged) && (avg_viol <= out->constraint_violation);
        out->used_average = use_avg;
        memcpy(x_out, use_avg ? x_avg : x, (size_t)n * sizeof(double));
        memcpy(y_out, y, (size_t)m * sizeof(double));
    }
done:
    out->status = status;
    free(col_sums); free(row_sums); free(tau); free(sigma); free(x); free(y);
    free(x_hat); free(x_prev); free(at_y); free(a_xhat); free(x_sum); free(x_avg); free(ax);
    return status;
}
