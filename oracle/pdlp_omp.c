/*
 * pdlp_omp.c -- the restarted, reflected Halpern PDHG of the GPU engine,
 * restated in FP64 with OpenMP over all host cores.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h): the multi-core CPU baseline of the new algorithm that
 * bench.py times beside the GPU, and a second, independent implementation
 * the tests compare the GPU engine's FP64 mode with.
 *
 * The reference has no such solver (its PDHG is src/first_order_solver.cpp:
 * 103-279, a plain diagonal-step loop).  The algorithm follows the PDLP line
 * of work (Applegate et al. 2021; Lu & Yang, cuPDLP 2023; the restarted
 * Halpern PDHG of cuPDLPx 2025):
 *   * Ruiz equilibration (10 sweeps of inf-norms) then Pock-Chambolle
 *     (alpha = 1, l1 norms), each sweep computing row and column norms from
 *     the scales of the previous sweep;
 *   * eta = 0.95 / ||A_hat||_2 from 40 rounds of power iteration, primal
 *     weight omega = ||c_hat|| / ||b_hat||, tau = eta / omega, sigma = eta omega;
 *   * step T(x, y): x+ = proj(x - tau (c + A'y)), y+ = proj(y + sigma (A(2x+ - x) - b)),
 *     then z <- w (2 T(z) - z) + (1 - w) z0, w = (k+1)/(k+2), A z carried;
 *   * every K steps the FP64 check of the unscaled problem for z and T(z),
 *     the KKT restart rule (beta 0.2 / 0.8 / 0.36), and the primal-weight
 *     update (theta 0.5) from the move of the anchor.
 * The rel-KKT is the one orc_kkt evaluates (PDLP's definition).
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#define NEW(T, k) ((T*)calloc((size_t)((k) > 0 ? (k) : 1), sizeof(T)))

static double now_s(void) { return omp_get_wtime(); }

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* out[i] = sum_k val[k] * x[idx[k]] over segment i (rows of CSR or columns of CSC) */
static void seg_spmv(int segs, const int* ptr, const int* idx, const double* val, const double* x, double* out) {
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < segs; ++i) {
        double s = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[idx[k]];
        out[i] = s;
    }
}

typedef struct {
    double eta, omega, tau, sigma, b_norm, c_norm, eps, eps_obj;
    double kkt_last_restart, kkt_prev_candidate;
    double rel_kkt[2], rel_primal[2], rel_dual[2], rel_gap[2], pobj[2], dobj[2], kkt_w[2];
    long long iter, iter_last_restart, max_iter;
    int done, status, restart, restart_to_cand, best, restarts;
} state_t;

/* Checkpoint of the iteration state at a check: the state record and the
 * carried vectors (scaled space).  The scaled problem itself is rebuilt by
 * the same deterministic set-up on resume.  Written to path.tmp, then renamed. */
static const char kCkptMagic[8] = {'B', 'O', 'P', 'D', 'L', 'P', 'C', '1'};

static int ckpt_write(const char* path, int m, int n, const state_t* st, double* const* xs, double* const* ys) {
    char tmp[4096];
    if (snprintf(tmp, sizeof(tmp), "%s.tmp", path) >= (int)sizeof(tmp)) return -1;
    FILE* f = fopen(tmp, "wb");
    if (!f) return -1;
    const long long dims[3] = {m, n, (long long)sizeof(state_t)};
    int ok = fwrite(kCkptMagic, 1, 8, f) == 8 && fwrite(dims, sizeof(dims), 1, f) == 1 &&
             fwrite(st, sizeof(state_t), 1, f) == 1;
    for (int a = 0; ok && a < 3; ++a) ok = fwrite(xs[a], sizeof(double), (size_t)n, f) == (size_t)n;
    for (int a = 0; ok && a < 5; ++a) ok = fwrite(ys[a], sizeof(double), (size_t)m, f) == (size_t)m;
    if (fclose(f) != 0) ok = 0;
    if (!ok || rename(tmp, path) != 0) {
        remove(tmp);
        return -1;
    }
    return 0;
}

static int ckpt_read(const char* path, int m, int n, state_t* st, double* const* xs, double* const* ys) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    char magic[8];
    long long dims[3];
    state_t s;
    int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kCkptMagic, 8) == 0 && fread(dims, sizeof(dims), 1, f) == 1 &&
             dims[0] == m && dims[1] == n && dims[2] == (long long)sizeof(state_t) && fread(&s, sizeof(s), 1, f) == 1;
    for (int a = 0; ok && a < 3; ++a) ok = fread(xs[a], sizeof(double), (size_t)n, f) == (size_t)n;
    for (int a = 0; ok && a < 5; ++a) ok = fread(ys[a], sizeof(double), (size_t)m, f) == (size_t)m;
    fclose(f);
    if (!ok) return -1;
    *st = s;
    return 0;
}

/* q: [0,1] c'x [2,3] ||viol||^2 [4,5] scaled [6,7] b'y [8,9] ||dual res||^2
 * [10,11] scaled [12,13] sum l plam+ + u plam-, for z [even] and T(z) [odd]. */
static void decide(state_t* st, const double* q) {
    const double w = st->omega;
    int bad = 0;
    for (int p = 0; p < 2; ++p) {
        const double pobj = q[0 + p], dobj = -q[6 + p] + q[12 + p];
        st->pobj[p] = pobj;
        st->dobj[p] = dobj;
        st->rel_primal[p] = sqrt(q[2 + p]) / (1.0 + st->b_norm);
        st->rel_dual[p] = sqrt(q[8 + p]) / (1.0 + st->c_norm);
        st->rel_gap[p] = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
        double k = st->rel_primal[p];
        if (st->rel_dual[p] > k) k = st->rel_dual[p];
        if (st->rel_gap[p] > k) k = st->rel_gap[p];
        st->rel_kkt[p] = k;
        const double g = pobj - dobj;
        st->kkt_w[p] = sqrt(w * q[4 + p] + q[10 + p] / w + g * g);
        if (!isfinite(k)) bad = 1;
    }
    const int best = st->rel_kkt[1] < st->rel_kkt[0] ? 1 : 0;
    st->best = best;
    if (bad) { st->done = 1; st->status = 2; return; }
    if (st->rel_kkt[best] <= st->eps &&
        (st->eps_obj <= 0.0 || fabs(st->pobj[best] - st->dobj[best]) <= st->eps_obj * (1.0 + fabs(st->pobj[best])))) {
        st->done = 1;
        st->status = 0;
        return;
    }
    if (st->iter >= st->max_iter) { st->done = 1; st->status = 1; return; }
    if (st->iter == 0) {
        st->kkt_last_restart = st->kkt_w[0];
        st->kkt_prev_candidate = st->kkt_w[0];
        return;
    }
    const int cand = st->kkt_w[1] < st->kkt_w[0] ? 1 : 0;
    const double kc = st->kkt_w[cand];
    const long long since = st->iter - st->iter_last_restart;
    const int restart = (kc <= 0.2 * st->kkt_last_restart) ||
                        (kc <= 0.8 * st->kkt_last_restart && kc > st->kkt_prev_candidate) ||
                        ((double)since >= 0.36 * (double)st->iter);
    st->kkt_prev_candidate = kc;
    if (restart) {
        st->restart = 1;
        st->restart_to_cand = cand;
        st->restarts += 1;
        st->kkt_last_restart = kc;
    }
}

int orc_pdlp_omp(int m, int n, const int* rp, const int* ci, const double* v, const double* b,
                 const double* c, const double* lb, const double* ub, const unsigned char* is_eq,
                 const orc_pdlp_opts* o, double* x_out, double* y_out, orc_pdlp_out* out) {
    const double t0 = now_s();
    const long long nnz = m > 0 ? rp[m] : 0;
    const int threads = o->threads > 0 ? o->threads : omp_get_max_threads();
    const int K = o->check_every > 0 ? o->check_every : 64;

    /* CSC of A with each column's rows ascending, and where each CSR entry went */
    int* cp = NEW(int, n + 1);
    int* ri = NEW(int, nnz);
    long long* perm = NEW(long long, nnz); /* CSR position -> CSC position */
    for (long long k = 0; k < nnz; ++k) cp[ci[k] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    {
        int* pos = NEW(int, n);
        memcpy(pos, cp, sizeof(int) * (size_t)n);
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int p = pos[ci[k]]++;
                ri[p] = i;
                perm[k] = p;
            }
        free(pos);
    }
    double* cv = NEW(double, nnz);   /* unscaled, CSC order */
    double* rvh = NEW(double, nnz);  /* scaled, CSR order */
    double* cvh = NEW(double, nnz);  /* scaled, CSC order */
    double *rs = NEW(double, m), *cs = NEW(double, n), *rnorm = NEW(double, m), *cnorm = NEW(double, n);
    double *ch = NEW(double, n), *lh = NEW(double, n), *uh = NEW(double, n), *bh = NEW(double, m);
    double *xk = NEW(double, n), *x0 = NEW(double, n), *xp = NEW(double, n), *atyv = NEW(double, n);
    double *yk = NEW(double, m), *y0 = NEW(double, m), *yp = NEW(double, m), *axv = NEW(double, m);
    double *ax = NEW(double, m), *ax0 = NEW(double, m), *ax_cand = NEW(double, m);
    double *xo_cur = NEW(double, n), *xo_cand = NEW(double, n), *yo_cur = NEW(double, m), *yo_cand = NEW(double, m);
    double *pw_v = NEW(double, n), *pw_u = NEW(double, m);
    for (long long k = 0; k < nnz; ++k) cv[perm[k]] = v[k];
    const double t_prep = now_s() - t0;

    state_t st;
    memset(&st, 0, sizeof(st));
    st.eps = o->eps_rel_kkt;
    st.eps_obj = o->eps_objective;
    st.max_iter = o->max_iterations;
    double b2 = 0.0, c2 = 0.0;
    for (int i = 0; i < m; ++i) b2 += b[i] * b[i];
    for (int j = 0; j < n; ++j) c2 += c[j] * c[j];
    st.b_norm = sqrt(b2);
    st.c_norm = sqrt(c2);
    double q[14];
    /* reduction targets are shared: declared outside the parallel region */
    double q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, q6 = 0, q7 = 0;
    double q8 = 0, q9 = 0, q10 = 0, q11 = 0, q12 = 0, q13 = 0;
    double chat2 = 0.0, bhat2 = 0.0, u2 = 0.0, w2 = 0.0, dx2 = 0.0, dy2 = 0.0;
    long long iters_run = 0;
    long long resumed_iter = 0;  /* iteration a checkpoint put the run at (0: a fresh start) */
    double t_setup = 0.0;

#pragma omp parallel num_threads(threads)
    {
        /* ---- scaling: Ruiz (inf norms) then Pock-Chambolle (l1 norms) ---- */
#pragma omp for
        for (int i = 0; i < m; ++i) rs[i] = 1.0;
#pragma omp for
        for (int j = 0; j < n; ++j) cs[j] = 1.0;
        for (int sweep = 0; sweep < o->ruiz_iterations + (o->pock_chambolle ? 1 : 0); ++sweep) {
            const int l1 = sweep >= o->ruiz_iterations;
#pragma omp for schedule(dynamic, 256)
            for (int i = 0; i < m; ++i) {
                double a = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    const double t = fabs(v[k] * cs[ci[k]]);
                    a = l1 ? a + t : (t > a ? t : a);
                }
                rnorm[i] = rs[i] * a;
            }
#pragma omp for schedule(dynamic, 256)
            for (int j = 0; j < n; ++j) {
                double a = 0.0;
                for (int k = cp[j]; k < cp[j + 1]; ++k) {
                    const double t = fabs(cv[k] * rs[ri[k]]);
                    a = l1 ? a + t : (t > a ? t : a);
                }
                cnorm[j] = cs[j] * a;
            }
#pragma omp for
            for (int i = 0; i < m; ++i)
                if (rnorm[i] > 0.0 && isfinite(rnorm[i])) rs[i] /= sqrt(rnorm[i]);
#pragma omp for
            for (int j = 0; j < n; ++j)
                if (cnorm[j] > 0.0 && isfinite(cnorm[j])) cs[j] /= sqrt(cnorm[j]);
        }
#pragma omp for
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) rvh[k] = rs[i] * v[k] * cs[ci[k]];
#pragma omp for
        for (int j = 0; j < n; ++j)
            for (int k = cp[j]; k < cp[j + 1]; ++k) cvh[k] = cs[j] * cv[k] * rs[ri[k]];
        /* scaled data and the starting point (projection of 0, or of x_start / y_start) */
#pragma omp for reduction(+ : chat2)
        for (int j = 0; j < n; ++j) {
            ch[j] = c[j] * cs[j];
            lh[j] = lb[j] / cs[j];
            uh[j] = ub[j] / cs[j];
            xk[j] = x0[j] = xp[j] = clampd(o->x_start ? o->x_start[j] / cs[j] : 0.0, lh[j], uh[j]);
            chat2 += ch[j] * ch[j];
        }
#pragma omp for reduction(+ : bhat2)
        for (int i = 0; i < m; ++i) {
            bh[i] = b[i] * rs[i];
            double ys = o->y_start ? o->y_start[i] / rs[i] : 0.0;
            if (!is_eq[i] && ys < 0.0) ys = 0.0;
            yk[i] = y0[i] = yp[i] = ys;
            bhat2 += bh[i] * bh[i];
        }
        /* power iteration for ||A_hat||_2 */
#pragma omp for
        for (int j = 0; j < n; ++j) {
            unsigned h = (unsigned)j * 2654435761u + 12345u;
            h ^= h >> 13;
            h *= 0x5bd1e995u;
            h ^= h >> 15;
            pw_v[j] = 0.5 + (double)(h & 0xffffu) / 65536.0;
        }
#pragma omp single
        w2 = 0.0;
#pragma omp for reduction(+ : w2)
        for (int j = 0; j < n; ++j) w2 += pw_v[j] * pw_v[j];
#pragma omp for
        for (int j = 0; j < n; ++j) pw_v[j] /= sqrt(w2);
        for (int it = 0; it < 40; ++it) {
            seg_spmv(m, rp, ci, rvh, pw_v, pw_u);
#pragma omp single
            u2 = w2 = 0.0;
#pragma omp for reduction(+ : u2)
            for (int i = 0; i < m; ++i) u2 += pw_u[i] * pw_u[i];
            seg_spmv(n, cp, ri, cvh, pw_u, pw_v);
#pragma omp for reduction(+ : w2)
            for (int j = 0; j < n; ++j) w2 += pw_v[j] * pw_v[j];
            if (w2 > 0.0) {
#pragma omp for
                for (int j = 0; j < n; ++j) pw_v[j] /= sqrt(w2);
            }
        }
#pragma omp single
        {
            const double smax = sqrt(u2), cn = sqrt(chat2), bn = sqrt(bhat2);
            st.eta = smax > 0.0 ? 0.95 / smax : 1.0;
            st.omega = o->primal_weight_start > 0.0 ? o->primal_weight_start
                                                    : ((cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0);
            st.tau = st.eta / st.omega;
            st.sigma = st.eta * st.omega;
            st.kkt_prev_candidate = INFINITY;
        }
        seg_spmv(m, rp, ci, rvh, xp, ax);
#pragma omp for
        for (int i = 0; i < m; ++i) ax0[i] = ax[i];
#pragma omp single
        {
            t_setup = now_s() - t0;
            if (o->checkpoint_path) {  /* continue a checkpointed run of this LP */
                double* const xs[3] = {xk, x0, xp};
                double* const ys[5] = {yk, y0, yp, ax, ax0};
                state_t saved;
                if (ckpt_read(o->checkpoint_path, m, n, &saved, xs, ys) == 0) {
                    st = saved;
                    st.max_iter = o->max_iterations;
                    st.eps = o->eps_rel_kkt;
                    st.eps_obj = o->eps_objective;
                    resumed_iter = st.iter;
                }
            }
        }

        for (;;) {
#pragma omp single
            if (o->checkpoint_path && o->checkpoint_every > 0 && st.iter > resumed_iter &&
                st.iter % o->checkpoint_every == 0) {
                double* const xs[3] = {xk, x0, xp};
                double* const ys[5] = {yk, y0, yp, ax, ax0};
                ckpt_write(o->checkpoint_path, m, n, &st, xs, ys);
            }
            /* ---- the FP64 check of z and T(z) on the unscaled problem ---- */
#pragma omp single
            q0 = q1 = q2 = q3 = q4 = q5 = q6 = q7 = q8 = q9 = q10 = q11 = q12 = q13 = 0.0;
#pragma omp for reduction(+ : q0, q1)
            for (int j = 0; j < n; ++j) {
                xo_cur[j] = clampd(cs[j] * xk[j], lb[j], ub[j]);
                xo_cand[j] = clampd(cs[j] * xp[j], lb[j], ub[j]);
                q0 += c[j] * xo_cur[j];
                q1 += c[j] * xo_cand[j];
            }
#pragma omp for schedule(dynamic, 256) reduction(+ : q2, q3, q4, q5, q6, q7)
            for (int i = 0; i < m; ++i) {
                double a0 = 0.0, a1 = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    a0 += v[k] * xo_cur[ci[k]];
                    a1 += v[k] * xo_cand[ci[k]];
                }
                const double r0 = a0 - b[i], r1 = a1 - b[i];
                const double v0 = is_eq[i] ? fabs(r0) : (r0 > 0.0 ? r0 : 0.0);
                const double v1 = is_eq[i] ? fabs(r1) : (r1 > 0.0 ? r1 : 0.0);
                q2 += v0 * v0;
                q3 += v1 * v1;
                q4 += (rs[i] * v0) * (rs[i] * v0);
                q5 += (rs[i] * v1) * (rs[i] * v1);
                ax_cand[i] = rs[i] * a1;
                double yc = rs[i] * yk[i];
                if (!is_eq[i] && yc < 0.0) yc = 0.0;
                yo_cur[i] = yc;
                yo_cand[i] = rs[i] * yp[i];
                q6 += b[i] * yc;
                q7 += b[i] * yo_cand[i];
            }
#pragma omp for schedule(dynamic, 256) reduction(+ : q8, q9, q10, q11, q12, q13)
            for (int j = 0; j < n; ++j) {
                double a0 = 0.0, a1 = 0.0;
                for (int k = cp[j]; k < cp[j + 1]; ++k) {
                    a0 += cv[k] * yo_cur[ri[k]];
                    a1 += cv[k] * yo_cand[ri[k]];
                }
                const int lf = isfinite(lb[j]), uf = isfinite(ub[j]);
                for (int p = 0; p < 2; ++p) {
                    const double lam = c[j] + (p ? a1 : a0);
                    double pl;
                    if (lf && uf) pl = lam;
                    else if (lf) pl = lam > 0.0 ? lam : 0.0;
                    else if (uf) pl = lam < 0.0 ? lam : 0.0;
                    else pl = 0.0;
                    const double d = lam - pl, sd = cs[j] * d;
                    const double t = pl > 0.0 ? lb[j] * pl : (pl < 0.0 ? ub[j] * pl : 0.0);
                    if (p) { q9 += d * d; q11 += sd * sd; q13 += t; }
                    else { q8 += d * d; q10 += sd * sd; q12 += t; }
                }
            }
#pragma omp single
            {
                q[0] = q0; q[1] = q1; q[2] = q2; q[3] = q3; q[4] = q4; q[5] = q5; q[6] = q6;
                q[7] = q7; q[8] = q8; q[9] = q9; q[10] = q10; q[11] = q11; q[12] = q12; q[13] = q13;
                decide(&st, q);
                dx2 = dy2 = 0.0;
            }
            if (st.done) break;
            if (st.restart) {
                const int to_cand = st.restart_to_cand;
#pragma omp for reduction(+ : dx2)
                for (int j = 0; j < n; ++j) {
                    const double xn = to_cand ? xp[j] : xk[j];
                    dx2 += (xn - x0[j]) * (xn - x0[j]);
                    xk[j] = x0[j] = xn;
                }
#pragma omp for reduction(+ : dy2)
                for (int i = 0; i < m; ++i) {
                    double yn = yk[i];
                    if (to_cand) {
                        yn = yp[i];
                        ax[i] = ax_cand[i];
                    }
                    dy2 += (yn - y0[i]) * (yn - y0[i]);
                    yk[i] = y0[i] = yn;
                    ax0[i] = ax[i];
                }
#pragma omp single
                {
                    const double dx = sqrt(dx2), dy = sqrt(dy2);
                    if (dx > 1e-10 && dy > 1e-10) {
                        st.omega = exp(0.5 * log(dy / dx) + 0.5 * log(st.omega));
                        st.tau = st.eta / st.omega;
                        st.sigma = st.eta * st.omega;
                    }
                    st.iter_last_restart = st.iter;
                    st.restart = 0;
                }
            }
            /* ---- K steps ---- */
            const long long left = st.max_iter - st.iter;
            const int k = (int)(left < K ? left : K);
            for (int t = 0; t < k; ++t) {
                const double kk = (double)(st.iter - st.iter_last_restart + t);
                const double w = (kk + 1.0) / (kk + 2.0);
                seg_spmv(n, cp, ri, cvh, yk, atyv);
#pragma omp for
                for (int j = 0; j < n; ++j) {
                    const double xo = xk[j];
                    const double xt = clampd(xo - st.tau * (ch[j] + atyv[j]), lh[j], uh[j]);
                    xp[j] = xt;
                    xk[j] = w * (2.0 * xt - xo) + (1.0 - w) * x0[j];
                }
                seg_spmv(m, rp, ci, rvh, xp, axv);
#pragma omp for
                for (int i = 0; i < m; ++i) {
                    const double axo = ax[i], yo = yk[i];
                    double yn = yo + st.sigma * (2.0 * axv[i] - axo - bh[i]);
                    if (!is_eq[i] && yn < 0.0) yn = 0.0;
                    yp[i] = yn;
                    yk[i] = w * (2.0 * yn - yo) + (1.0 - w) * y0[i];
                    ax[i] = w * (2.0 * axv[i] - axo) + (1.0 - w) * ax0[i];
                }
            }
#pragma omp single
            {
                st.iter += k;
                iters_run += k;
            }
        }
    }
    const int bst = st.best;
    if (x_out) memcpy(x_out, bst ? xo_cand : xo_cur, sizeof(double) * (size_t)n);
    if (y_out) memcpy(y_out, bst ? yo_cand : yo_cur, sizeof(double) * (size_t)m);
    out->status = st.status;
    out->iterations = st.iter;
    out->restarts = st.restarts;
    out->rel_kkt = st.rel_kkt[bst];
    out->rel_primal = st.rel_primal[bst];
    out->rel_dual = st.rel_dual[bst];
    out->rel_gap = st.rel_gap[bst];
    out->primal_obj = st.pobj[bst];
    out->dual_obj = st.dobj[bst];
    out->threads = threads;
    out->step_size = st.eta;
    out->primal_weight = st.omega;
    out->prep_s = t_prep;
    out->setup_s = t_setup;
    out->solve_s = now_s() - t0;
    void* bufs[] = {cp, ri, perm, cv, rvh, cvh, rs, cs, rnorm, cnorm, ch, lh, uh, bh, xk, x0, xp, atyv, yk, y0, yp,
                    axv, ax, ax0, ax_cand, xo_cur, xo_cand, yo_cur, yo_cand, pw_v, pw_u};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i) free(bufs[i]);
    return st.status;
}
