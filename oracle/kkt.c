/*
 * kkt.c -- FP64 relative KKT error of a standard-form LP.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * Problem (reference sign convention, src/first_order_solver.cpp:68-99):
 *     min c'x  s.t.  A_E x = b_E,  A_I x <= b_I,  l <= x <= u
 * Lagrangian c'x + y'(Ax - b) with y_I >= 0, reduced cost lambda = c + A'y.
 * Dual function g(y) = -b'y + sum_j min_{l_j<=x_j<=u_j} lambda_j x_j
 *                    = -b'y + sum_j (l_j lambda_j^+ - u_j lambda_j^-)
 * evaluated at the projection of lambda onto the set where that minimum is
 * finite (lambda_j >= 0 if only l_j is finite, <= 0 if only u_j, 0 if neither).
 * The reference has no KKT measure; this follows PDLP (Applegate et al. 2021).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

void orc_kkt(int m, int n, const int* rp, const int* ci, const double* v,
             const double* b, const double* c, const double* lb, const double* ub,
             const unsigned char* is_eq, const double* x, const double* y,
             orc_kkt_out* out) {
    double* lam = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double rp2 = 0.0, rd2 = 0.0, bty = 0.0, b2 = 0.0, c2 = 0.0, pobj = 0.0, dterm = 0.0;
    double maxrow = 0.0, maxbnd = 0.0;

    for (int j = 0; j < n; ++j) lam[j] = c[j];
    for (int i = 0; i < m; ++i) {
        double ax = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            ax += v[k] * x[ci[k]];
            lam[ci[k]] += v[k] * y[i];
        }
        double r = ax - b[i];
        double viol = is_eq[i] ? fabs(r) : (r > 0.0 ? r : 0.0);
        rp2 += viol * viol;
        if (viol > maxrow) maxrow = viol;
        if (!is_eq[i] && y[i] < 0.0) rd2 += y[i] * y[i];
        bty += b[i] * y[i];
        b2 += b[i] * b[i];
    }
    for (int j = 0; j < n; ++j) {
        const int lf = isfinite(lb[j]), uf = isfinite(ub[j]);
        double l = lam[j], pl;
        if (lf && uf) pl = l;
        else if (lf) pl = l > 0.0 ? l : 0.0;
        else if (uf) pl = l < 0.0 ? l : 0.0;
        else pl = 0.0;
        rd2 += (l - pl) * (l - pl);
        if (pl > 0.0) dterm += lb[j] * pl;
        else if (pl < 0.0) dterm += ub[j] * pl;
        pobj += c[j] * x[j];
        c2 += c[j] * c[j];
        if (x[j] < lb[j] && lb[j] - x[j] > maxbnd) maxbnd = lb[j] - x[j];
        if (x[j] > ub[j] && x[j] - ub[j] > maxbnd) maxbnd = x[j] - ub[j];
    }
    out->r_primal = sqrt(rp2);
    out->r_dual = sqrt(rd2);
    out->primal_obj = pobj;
    out->dual_obj = -bty + dterm;
    out->gap_rel = fabs(pobj - out->dual_obj) / (1.0 + fabs(pobj) + fabs(out->dual_obj));
    out->b_norm = sqrt(b2);
    out->c_norm = sqrt(c2);
    out->rel_primal = out->r_primal / (1.0 + out->b_norm);
    out->rel_dual = out->r_dual / (1.0 + out->c_norm);
    double k = out->rel_primal;
    if (out->rel_dual > k) k = out->rel_dual;
    if (out->gap_rel > k) k = out->gap_rel;
    out->rel_kkt = k;
    out->max_row_viol = maxrow;
    out->max_bound_viol = maxbnd;
    free(lam);
}
