// pdlp_core.hpp -- per-LP state and step logic of the restarted, reflected
// Halpern PDHG, shared by the single-LP engine (pdlp.hip) and the batched
// engine (batch.hip).  Everything here acts on one LP's PdlpState; the
// batched engine keeps one state per LP and maps every row and column to
// its LP.
#pragma once

#include "common.hpp"

namespace bo {

struct PdlpState {
    double eta, omega, tau, sigma;
    double b_norm, c_norm, eps;
    double eps_obj;     // > 0: also stop only when |c'x - dual obj| <= eps_obj (1 + |c'x|)
    double kkt_last_restart, kkt_prev_candidate;
    // [0] current iterate z, [1] candidate T(z)
    double rel_kkt[2], rel_primal[2], rel_dual[2], rel_gap[2], pobj[2], dobj[2], kkt_w[2];
    double sigma_max;
    double omega_init;  // > 0: starting primal weight (a warm start's), else ||c_hat|| / ||b_hat||
    double best_kkt;    // lowest rel-KKT seen, updated when it halves
    long long best_kkt_iter;  // iteration of that update
    long long stall_window;   // > 0: stop as STALLED after max(this, iter / 2) iterations without a halving
    long long iter, iter_last_restart, max_iter;
    int done, status, restart, restart_to_cand, best, restarts, checks, pad;
};

// Halpern weight of step t of the current interval, from the state's
// iteration counters.
__device__ inline double halpern_w(long long iter, long long iter_last_restart, int t) {
    const double k = static_cast<double>(iter - iter_last_restart + t);
    return (k + 1.0) / (k + 2.0);
}
__device__ inline double halpern_w(const PdlpState* st, int t) {
    return halpern_w(st->iter, st->iter_last_restart, t);
}

// The step's updates with every fused multiply-add written out.  Left to the
// compiler, contraction follows instruction scheduling, so a change elsewhere
// in a kernel (where an operand is loaded, say) can change which product is
// rounded; a solve with FP32 iterates is sensitive to one such rounding (it
// once moved the 200,000 x 500,000 Zipf LP from 11,008 to 12,288 iterations).
// Every kernel that takes a step uses these.
// x - tau (c + A'y), before the projection
__device__ inline double primal_move(double xo, double tau, double c, double aty) {
    return fma(-tau, c + aty, xo);
}
// y + sigma (2 A x+ - A x - b), before the projection
__device__ inline double dual_move(double yo, double sigma, double ax_plus, double axo, double b) {
    return fma(sigma, (2.0 * ax_plus - axo) - b, yo);
}
// Halpern updates w (2 t - z) + (1 - w) z0 of x, of y and of A x
__device__ inline double halpern_x(double w, double xt, double xo, double x0) {
    return fma(x0, 1.0 - w, w * fma(xt, 2.0, -xo));
}
__device__ inline double halpern_y(double w, double yt, double yo, double y0) {
    return fma(w, fma(yt, 2.0, -yo), (1.0 - w) * y0);
}
__device__ inline double halpern_ax(double w, double ax_plus, double axo, double ax0) {
    return fma(w, 2.0 * ax_plus - axo, (1.0 - w) * ax0);
}

template <class T>
__host__ __device__ inline T clampT(T v, T lo, T hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Column j: primal step from A'y (gathered y copy), Halpern update of x.
// inst (batched engine): the LP of every column, whose state st[inst[j]]
// drives the step; a finished LP is left untouched.
template <class T>
struct EpiPrimal {
    using TXv = T;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const T* yg;       // gathered: the current dual iterate, rounded to T
    const PdlpState* st;
    const T* c;
    const T* lb;
    const T* ub;
    double* xk;        // Halpern iterate (in/out)
    const double* x0;  // restart anchor
    T* xp;             // x+ (gathered by the CSR pass)
    int t;             // step index within the interval
    const int* inst;   // nullptr: one LP
    const unsigned char* wg_live = nullptr;  // batched: 0 for a reduce workgroup whose LPs all stopped
    __device__ const T* vec(int) const { return yg; }
    __device__ bool skip(int j) const { return inst && st[inst[j]].done; }
    __device__ bool wg_dead(int b) const { return wg_live && !wg_live[b]; }
    // Column data the step reads besides A'y, loaded by a reduction kernel
    // before the partials so the two sets of loads overlap.  The state's
    // step size, iteration counters and done flag are part of it, loaded raw
    // and used only in finish: read there, or turned into the Halpern weight
    // up front, they would put a round trip before the stores or before the
    // partial loads.  The per-lane state pointer keeps them vector loads,
    // whose waits are in order (a scalar-load wait waits for every scalar
    // load in flight).  A step of a finished LP stores nothing.
    struct Pre {
        double xo, x0, tau;
        long long it, it_restart;
        T c, lb, ub;
        int stop;  // the column's LP has finished
    };
    __device__ Pre load(int j) const {
        const PdlpState* s = inst ? st + inst[j] : st;
        return Pre{xk[j], x0[j], s->tau, s->iter, s->iter_last_restart, c[j], lb[j], ub[j], s->done};
    }
    __device__ void finish(int j, double acc, const Pre& p) const {
        if (p.stop) return;
        const T xt = clampT(static_cast<T>(primal_move(p.xo, p.tau, static_cast<double>(p.c), acc)), p.lb, p.ub);
        xp[j] = xt;
        xk[j] = halpern_x(halpern_w(p.it, p.it_restart, t), static_cast<double>(xt), p.xo, p.x0);
    }
    __device__ void operator()(int j, const double* acc, double*) const {
        if (inst && st[inst[j]].done) return;
        finish(j, acc[0], load(j));
    }
};

// Row i: dual step from A x+ (gathered x+), Halpern update of y and of A x.
template <class T>
struct EpiDual {
    using TXv = T;
    static constexpr int NV = 1;
    static constexpr int Q = 0;
    const T* xp;       // gathered: x+
    const PdlpState* st;
    const T* b;
    const unsigned char* eq;
    double* ax;        // A z (carried, in/out)
    const double* ax0; // A z0
    double* yk;        // Halpern iterate (in/out)
    const double* y0;  // restart anchor
    T* yp;             // y+ (restart candidate)
    T* yg;             // yk rounded to T (gathered by the next CSC pass)
    int t;
    const int* inst;   // nullptr: one LP
    const unsigned char* wg_live = nullptr;  // batched: 0 for a reduce workgroup whose LPs all stopped
    __device__ const T* vec(int) const { return xp; }
    __device__ bool skip(int i) const { return inst && st[inst[i]].done; }
    __device__ bool wg_dead(int b) const { return wg_live && !wg_live[b]; }
    // Row data the step reads besides A x+, with the state's step size and
    // iteration counters (see EpiPrimal::Pre).
    struct Pre {
        double axo, yo, y0, ax0, sigma;
        long long it, it_restart;
        T b;
        unsigned char eq;
        int stop;  // the row's LP has finished
    };
    __device__ Pre load(int i) const {
        const PdlpState* s = inst ? st + inst[i] : st;
        return Pre{ax[i], yk[i], y0[i], ax0[i], s->sigma, s->iter, s->iter_last_restart, b[i], eq[i], s->done};
    }
    __device__ void finish(int i, double acc, const Pre& p) const {
        if (p.stop) return;
        double yn = dual_move(p.yo, p.sigma, acc, p.axo, static_cast<double>(p.b));
        if (!p.eq && yn < 0.0) yn = 0.0;
        const T yt = static_cast<T>(yn);
        yp[i] = yt;
        const double w = halpern_w(p.it, p.it_restart, t);
        const double ynew = halpern_y(w, static_cast<double>(yt), p.yo, p.y0);
        yk[i] = ynew;
        yg[i] = static_cast<T>(ynew);
        ax[i] = halpern_ax(w, acc, p.axo, p.ax0);
    }
    __device__ void operator()(int i, const double* acc, double*) const {
        if (inst && st[inst[i]].done) return;
        finish(i, acc[0], load(i));
    }
};

// eta, omega, tau, sigma from the power-iteration estimate ||A_hat u||^2 and
// ||c_hat||^2, ||b_hat||^2; the iteration counters start over.
__device__ inline void init_state_core(PdlpState* st, double pw_u2, double c2, double b2) {
    const double smax = sqrt(pw_u2);
    st->sigma_max = smax;
    st->eta = smax > 0.0 ? 0.95 / smax : 1.0;
    const double cn = sqrt(c2), bn = sqrt(b2);
    st->omega = st->omega_init > 0.0 ? st->omega_init : ((cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0);
    st->tau = st->eta / st->omega;
    st->sigma = st->eta * st->omega;
    st->iter = 0;
    st->iter_last_restart = 0;
    st->done = 0;
    st->status = 1;
    st->restart = 0;
    st->restart_to_cand = 0;
    st->best = 0;
    st->restarts = 0;
    st->checks = 0;
    st->kkt_last_restart = 0.0;
    st->kkt_prev_candidate = INFINITY;
    st->best_kkt = INFINITY;
    st->best_kkt_iter = 0;
}

// The check's sums q (FP64, unscaled problem):
//   [0,1] c'x  [2,3] ||primal violation||^2  [4,5] the same, scaled  [6,7] b'y
//   [8,9] ||dual residual||^2  [10,11] the same, scaled  [12,13] sum l plam+ + u plam-
// for the current iterate [even] and the candidate T(z) [odd]; then the
// stopping test (rel-KKT <= eps and, when eps_obj > 0, the objective gap
// |c'x - dual obj| <= eps_obj (1 + |c'x|): with x and y feasible to the KKT
// tolerance the optimum lies between the two objectives, so c'x is then
// within eps_obj of it relative to 1 + |c'x|), the stall test (no halving of the best rel-KKT within
// max(stall_window, iter / 2) iterations: restart intervals grow with the
// iteration count, and so do plateaus of a converging run), the KKT restart
// rule (beta 0.2 / 0.8 / 0.36) and the bookkeeping.
__device__ inline void decide_core(PdlpState* st, const double* q) {
    const double w = st->omega;
    bool bad = false;
    for (int p = 0; p < 2; ++p) {
        const double pobj = q[0 + p];
        const double dobj = -q[6 + p] + q[12 + p];
        const double rp = sqrt(q[2 + p]), rd = sqrt(q[8 + p]);
        st->pobj[p] = pobj;
        st->dobj[p] = dobj;
        st->rel_primal[p] = rp / (1.0 + st->b_norm);
        st->rel_dual[p] = rd / (1.0 + st->c_norm);
        st->rel_gap[p] = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
        double k = st->rel_primal[p];
        if (st->rel_dual[p] > k) k = st->rel_dual[p];
        if (st->rel_gap[p] > k) k = st->rel_gap[p];
        st->rel_kkt[p] = k;
        const double g = pobj - dobj;
        st->kkt_w[p] = sqrt(w * q[4 + p] + q[10 + p] / w + g * g);
        if (!isfinite(k)) bad = true;
    }
    st->checks += 1;
    const int best = st->rel_kkt[1] < st->rel_kkt[0] ? 1 : 0;
    st->best = best;
    if (bad) {
        st->done = 1;
        st->status = 2;
        return;
    }
    if (st->rel_kkt[best] <= st->eps &&
        (st->eps_obj <= 0.0 || fabs(st->pobj[best] - st->dobj[best]) <= st->eps_obj * (1.0 + fabs(st->pobj[best])))) {
        st->done = 1;
        st->status = 0;
        return;
    }
    if (st->iter >= st->max_iter) {
        st->done = 1;
        st->status = 1;
        return;
    }
    if (st->rel_kkt[best] < 0.5 * st->best_kkt) {
        st->best_kkt = st->rel_kkt[best];
        st->best_kkt_iter = st->iter;
    } else if (st->stall_window > 0 &&
               st->iter - st->best_kkt_iter >= (st->stall_window > st->iter / 2 ? st->stall_window : st->iter / 2)) {
        st->done = 1;
        st->status = 4;
        return;
    }
    if (st->iter == 0) {  // starting point
        st->kkt_last_restart = st->kkt_w[0];
        st->kkt_prev_candidate = st->kkt_w[0];
        return;
    }
    const int cand = st->kkt_w[1] < st->kkt_w[0] ? 1 : 0;
    const double kc = st->kkt_w[cand];
    const long long since = st->iter - st->iter_last_restart;
    const bool restart = (kc <= 0.2 * st->kkt_last_restart) ||
                         (kc <= 0.8 * st->kkt_last_restart && kc > st->kkt_prev_candidate) ||
                         (static_cast<double>(since) >= 0.36 * static_cast<double>(st->iter));
    st->kkt_prev_candidate = kc;
    if (restart) {
        st->restart = 1;
        st->restart_to_cand = cand;
        st->restarts += 1;
        st->kkt_last_restart = kc;
    }
}

// After a restart: primal-weight update (theta = 0.5) from the anchor move.
__device__ inline void restart_finish_core(PdlpState* st, double dx2, double dy2) {
    const double dx = sqrt(dx2), dy = sqrt(dy2);
    if (dx > 1e-10 && dy > 1e-10) {
        st->omega = exp(0.5 * log(dy / dx) + 0.5 * log(st->omega));
        st->tau = st->eta / st->omega;
        st->sigma = st->eta * st->omega;
    }
    st->iter_last_restart = st->iter;
    st->restart = 0;
}

}  // namespace bo
