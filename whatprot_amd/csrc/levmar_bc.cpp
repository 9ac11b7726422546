// Box-constrained Levenberg-Marquardt for the detach-rate M-step.
//
// least_squares_fit_of_offset_exponential (util/levmar-helper.cc:44-85) fits
// y_t ~ base + initial * exp(-decay * t) with weights w_t by calling levmar
// 2.6's dlevmar_bc_der (M. I. A. Lourakis; GPL, vendored by the reference as
// cc_code/extern/levmar-2.6, built without LAPACK) with bounds [0, 1] (a held
// parameter has lb = ub = its value), 10 iterations and default options. This
// file restates that published method so the fitted rates match:
//
//  * damped Gauss-Newton steps on the sqrt(w)-weighted residuals, projected
//    on the box, accepted when they cut ||e||^2 by the factor gamma, with
//    Nielsen's damping update (Kanzow, Yamashita, Fukushima, "Levenberg-
//    Marquardt methods with strong local convergence properties for solving
//    nonlinear equations with convex constraints", J. Comput. Appl. Math.
//    172, 2004, with Kanzow's starting damping mu0 = 0.5 tau ||e||^2);
//  * otherwise a backtracking line search along the LM direction when it is a
//    descent direction (Dennis & Schnabel, "Numerical Methods for
//    Unconstrained Optimization", algorithm A6.3.1, with projection), else a
//    projected-gradient Armijo search;
//  * the normal equations solved by Crout LU with implicit partial pivoting.
//
// Two arithmetic details of levmar 2.6 are kept because they change the last
// bits of results: its squared-norm routine sums in four interleaved
// accumulators over blocks of eight taken from the top, and its LU solver
// replaces a zero pivot with DBL_EPSILON.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "wp_internal.h"

namespace wp {
namespace {

constexpr int kM = 3;  // base, initial, initial_decay

struct Problem {
    int n;
    const double* sw;  // sqrt of the weights
    const double* x;   // sqrt(w) * y
    const double* lb;
    const double* ub;
};

// sqrt(w_t) * (base + initial * exp(-decay * t)) (levmar-helper.cc:23-29).
void model(const Problem& pr, const double* p, double* hx) {
    for (int t = 0; t < pr.n; t++) hx[t] = pr.sw[t] * (p[0] + p[1] * exp(-p[2] * t));
}

// Row-major n x 3 Jacobian (levmar-helper.cc:31-42).
void jacobian(const Problem& pr, const double* p, double* J) {
    for (int t = 0; t < pr.n; t++) {
        double g = exp(-p[2] * t);
        J[t * kM + 0] = pr.sw[t];
        J[t * kM + 1] = pr.sw[t] * g;
        J[t * kM + 2] = pr.sw[t] * -p[1] * t * g;
    }
}

// e = x - y and sum e^2, in levmar 2.6's summation order (see the header).
double residual_sq(double* e, const double* x, const double* y, int n) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int full = n & ~7;
    for (int top = full - 1; top > 0; top -= 8) {
        // element top - k goes to accumulator k mod 4
        for (int k = 0; k < 8; k++) {
            int i = top - k;
            e[i] = x[i] - y[i];
            s[k & 3] += e[i] * e[i];
        }
    }
    static const int tail_acc[7] = {0, 1, 2, 3, 0, 1, 2};
    int rem = n - full;
    for (int k = 0; k < rem; k++) {
        int i = full + k;
        e[i] = x[i] - y[i];
        s[tail_acc[7 - rem + k]] += e[i] * e[i];
    }
    return s[0] + s[1] + s[2] + s[3];
}

// Overflow-safe Euclidean norm (Blue's scaling), as levmar uses without BLAS.
double safe_norm(const double* v, int n) {
    double mx = 0.0;
    for (int i = n - 1; i >= 0; i--) {
        if (v[i] > mx) mx = v[i];
        else if (v[i] < -mx) mx = -v[i];
    }
    double s = 0.0;
    for (int i = n - 1; i >= 0; i--) {
        double q = v[i] / mx;
        s += q * q;
    }
    return mx * sqrt(s);
}

void project(double* p, const double* lb, const double* ub) {
    for (int i = kM - 1; i >= 0; i--) {
        double a = lb[i], b = p[i], c = ub[i];
        // median of (lb, p, ub)
        if (a >= b) p[i] = c >= a ? a : (c <= b ? b : c);
        else p[i] = c >= b ? b : (c <= a ? a : c);
    }
}

// Solves A x = b (3 x 3, row-major) by Crout LU with implicit pivoting;
// returns false when a row of A is zero.
bool lu_solve(const double* A, const double* b, double* x) {
    double a[kM * kM], scale[kM];
    int perm[kM];
    memcpy(a, A, sizeof(a));
    memcpy(x, b, sizeof(double) * kM);
    for (int i = 0; i < kM; i++) {
        double mx = 0.0;
        for (int j = 0; j < kM; j++) mx = fmax(mx, fabs(a[i * kM + j]));
        if (mx == 0.0) return false;
        scale[i] = 1.0 / mx;
    }
    for (int j = 0; j < kM; j++) {
        for (int i = 0; i < j; i++) {
            double s = a[i * kM + j];
            for (int k = 0; k < i; k++) s -= a[i * kM + k] * a[k * kM + j];
            a[i * kM + j] = s;
        }
        double best = 0.0;
        int piv = -1;
        for (int i = j; i < kM; i++) {
            double s = a[i * kM + j];
            for (int k = 0; k < j; k++) s -= a[i * kM + k] * a[k * kM + j];
            a[i * kM + j] = s;
            double q = scale[i] * fabs(s);
            if (q >= best) {  // later rows win ties
                best = q;
                piv = i;
            }
        }
        if (piv != j) {
            for (int k = 0; k < kM; k++) {
                double tmp = a[piv * kM + k];
                a[piv * kM + k] = a[j * kM + k];
                a[j * kM + k] = tmp;
            }
            scale[piv] = scale[j];
        }
        perm[j] = piv;
        if (a[j * kM + j] == 0.0) a[j * kM + j] = DBL_EPSILON;
        if (j != kM - 1) {
            double inv = 1.0 / a[j * kM + j];
            for (int i = j + 1; i < kM; i++) a[i * kM + j] *= inv;
        }
    }
    // forward substitution, skipping the leading zeros of the permuted b
    int first = 0;
    for (int i = 0; i < kM; i++) {
        int pi = perm[i];
        double s = x[pi];
        x[pi] = x[i];
        if (first != 0) {
            for (int j = first - 1; j < i; j++) s -= a[i * kM + j] * x[j];
        } else if (s != 0.0) {
            first = i + 1;
        }
        x[i] = s;
    }
    for (int i = kM - 1; i >= 0; i--) {
        double s = x[i];
        for (int j = i + 1; j < kM; j++) s -= a[i * kM + j] * x[j];
        x[i] = s / a[i * kM + i];
    }
    return true;
}

// Backtracking line search along dir from p (Dennis & Schnabel A6.3.1 with
// box projection). f0 = ||e(p)||^2. On success (return 0) q holds the new
// point, r its residual and *fq its ||e||^2. dir may be shortened in place.
int line_search(const Problem& pr, const double* p, double f0, const double* g, double* dir,
                double* q, double* r, double* fq, double stepmx, double steptl) {
    const double alpha = 1e-4;
    double half_f = 0.5 * f0;
    double len2 = 0.0;
    for (int i = kM - 1; i >= 0; i--) len2 += dir[i] * dir[i];
    double len = sqrt(len2);
    if (len > stepmx) {
        double s = stepmx / len;
        for (int i = kM - 1; i >= 0; i--) dir[i] *= s;
    }
    double slope = 0.0, rel = 0.0;
    for (int i = kM - 1; i >= 0; i--) {
        slope += g[i] * dir[i];
        double ri = fabs(dir[i]) / fmax(fabs(p[i]), 1.0);
        if (rel < ri) rel = ri;
    }
    double lambda_min = steptl / rel;
    double lambda = 1.0, prev_lambda = 0.0, prev_f = 0.0;
    bool first_back = true;
    for (int it = 150; it-- > 0;) {
        for (int i = kM - 1; i >= 0; i--) q[i] = p[i] + lambda * dir[i];
        project(q, pr.lb, pr.ub);
        model(pr, q, r);
        double ss = residual_sq(r, pr.x, r, pr.n);
        double f = 0.5 * ss;
        *fq = ss;
        if (f <= half_f + slope * alpha * lambda) return 0;
        if (lambda < lambda_min) return 1;
        if (!std::isfinite(f)) {
            lambda *= 0.1;
            first_back = true;
            continue;
        }
        double next;
        if (first_back) {  // quadratic model of f along the step
            next = -lambda * slope / ((f - half_f - slope) * 2.0);
            first_back = false;
        } else {  // cubic through the last two trial points
            double t1 = f - half_f - lambda * slope;
            double t2 = prev_f - half_f - prev_lambda * slope;
            double t3 = 1.0 / (lambda - prev_lambda);
            double a3 = 3.0 * t3 * (t1 / (lambda * lambda) - t2 / (prev_lambda * prev_lambda));
            double b = t3 * (t2 * lambda / (prev_lambda * prev_lambda) - t1 * prev_lambda / (lambda * lambda));
            double disc = b * b - a3 * slope;
            if (disc > b * b) next = (-b + (a3 < 0 ? -sqrt(disc) : sqrt(disc))) / a3;
            else next = (-b + (a3 < 0 ? sqrt(disc) : -sqrt(disc))) / a3;
            if (next > lambda * 0.5) next = lambda * 0.5;
        }
        prev_lambda = lambda;
        prev_f = f;
        lambda = next < lambda * 0.1 ? lambda * 0.1 : next;
    }
    return 1;
}

}  // namespace

void offset_exponential_fit(int n, const double* y, const double* w, const int* hold, double* p, int itmax) {
    const double tau = 1e-3, eps1 = 1e-17, eps2 = 1e-17, eps3 = 1e-17;
    const double eps2_sq = eps2 * eps2;
    const double gamma = 0.99995, rho = 1e-8, beta = 0.9, alpha = 1e-4;
    const double singular_eps = 1e-12;
    if (n < kM) return;  // fewer measurements than unknowns: levmar refuses
    double lb[kM], ub[kM];
    for (int i = 0; i < kM; i++) {
        lb[i] = hold[i] ? p[i] : 0.0;
        ub[i] = hold[i] ? p[i] : 1.0;
    }
    double* buf = new double[(size_t)n * (4 + kM)];
    double* sw = buf;
    double* x = sw + n;
    double* e = x + n;
    double* hx = e + n;
    double* J = hx + n;
    for (int t = 0; t < n; t++) sw[t] = sqrt(w[t]);
    for (int t = 0; t < n; t++) x[t] = sw[t] * y[t];
    Problem pr{n, sw, x, lb, ub};

    project(p, lb, ub);
    model(pr, p, hx);
    double err = residual_sq(e, x, hx, n);
    bool stop = !std::isfinite(err);
    double mu = 0.0, t_pg = 0.0;
    int nu = 2;
    bool grad_prev = false;
    double JtJ[kM * kM], diag[kM], g[kM], dp[kM], q[kM];
    for (int k = 0; k < itmax && !stop; k++) {
        if (err <= eps3) break;
        jacobian(pr, p, J);
        for (int i = 0; i < kM * kM; i++) JtJ[i] = 0.0;
        for (int i = 0; i < kM; i++) g[i] = 0.0;
        for (int l = n - 1; l >= 0; l--) {
            const double* row = &J[l * kM];
            for (int i = kM - 1; i >= 0; i--) {
                double a = row[i];
                for (int j = i; j >= 0; j--) JtJ[i * kM + j] += row[j] * a;
                g[i] += a * e[l];
            }
        }
        for (int i = kM - 1; i >= 0; i--)
            for (int j = i + 1; j < kM; j++) JtJ[i * kM + j] = JtJ[j * kM + i];
        // gradient norm over free parameters; a bound parameter counts as
        // converged when the gradient pushes it into its bound
        int active = 0, pinned = 0;
        double g_inf = 0.0, p_sq = 0.0;
        for (int i = 0; i < kM; i++) {
            if (p[i] == ub[i]) {
                active++;
                if (g[i] > 0.0) pinned++;
            } else if (p[i] == lb[i]) {
                active++;
                if (g[i] < 0.0) pinned++;
            } else if (g_inf < fabs(g[i])) {
                g_inf = fabs(g[i]);
            }
            diag[i] = JtJ[i * kM + i];
            p_sq += p[i] * p[i];
        }
        if (pinned == active && g_inf <= eps1) break;
        if (k == 0) mu = 0.5 * tau * err;

        while (true) {
            for (int i = 0; i < kM; i++) JtJ[i * kM + i] += mu;
            double dp_sq, err_new = 0.0;
            if (!lu_solve(JtJ, g, dp)) {
                mu *= nu;
                int nu2 = nu << 1;
                if (nu2 <= nu) {
                    stop = true;
                    break;
                }
                nu = nu2;
                for (int i = 0; i < kM; i++) JtJ[i * kM + i] = diag[i];
                continue;
            }
            for (int i = 0; i < kM; i++) q[i] = p[i] + dp[i];
            project(q, lb, ub);
            dp_sq = 0.0;
            for (int i = 0; i < kM; i++) {
                dp[i] = q[i] - p[i];
                dp_sq += dp[i] * dp[i];
            }
            if (dp_sq <= eps2_sq * p_sq) {
                stop = true;
                break;
            }
            if (dp_sq >= (p_sq + eps2) / (singular_eps * singular_eps)) {
                stop = true;
                break;
            }
            model(pr, q, hx);
            err_new = residual_sq(hx, x, hx, n);
            if (!std::isfinite(err_new) && !std::isfinite(safe_norm(hx, n))) {
                stop = true;
                break;
            }
            if (err_new <= gamma * err) {  // LM step accepted
                double dL = 0.0;
                for (int i = 0; i < kM; i++) dL += dp[i] * (mu * dp[i] + g[i]);
                if (dL > 0.0) {
                    double dF = err - err_new;
                    double r = 2.0 * dF / dL - 1.0;
                    r = 1.0 - r * r * r;
                    mu = mu * (r >= 0.3333333334 ? r : 0.3333333334);
                } else {
                    double cap = 0.1 * err_new;
                    mu = mu >= cap ? cap : mu;
                }
                nu = 2;
                for (int i = 0; i < kM; i++) p[i] = q[i];
                memcpy(e, hx, sizeof(double) * n);
                err = err_new;
                grad_prev = false;
                break;
            }
            // LM step rejected: search along it if it descends, else along
            // the projected negative gradient
            double slope = 0.0;
            for (int i = 0; i < kM; i++) {
                g[i] = -g[i];
                slope += g[i] * dp[i];
            }
            bool searched = false;
            if (slope <= -rho * pow(dp_sq, 2.1 / 2.0)) {
                double stepmx = 1e3 * fmax(sqrt(p_sq), 1.0);
                double steptl = 1e3 * sqrt(DBL_EPSILON);
                int rc = line_search(pr, p, err, g, dp, q, hx, &err_new, stepmx, steptl);
                searched = rc == 0 && std::isfinite(err_new);
                if (searched) grad_prev = false;
            }
            if (!searched) {
                double gn = 0.0;
                for (int i = 0; i < kM; i++) gn += g[i] * g[i];
                gn = sqrt(gn);
                double t0 = fmin(100.0 / (1.0 + gn), 1.0);
                bool found = false;
                for (t_pg = grad_prev ? t_pg : t0; t_pg > 1e-18; t_pg *= beta) {
                    for (int i = 0; i < kM; i++) q[i] = p[i] - t_pg * g[i];
                    project(q, lb, ub);
                    for (int i = 0; i < kM; i++) dp[i] = q[i] - p[i];
                    model(pr, q, hx);
                    err_new = residual_sq(hx, x, hx, n);
                    if (!std::isfinite(err_new) && !std::isfinite(safe_norm(hx, n))) {
                        stop = true;
                        break;
                    }
                    double gd = 0.0;
                    for (int i = 0; i < kM; i++) gd += g[i] * dp[i];
                    if (grad_prev && err_new <= err + 2.0 * 0.99999 * gd) {  // start was too short
                        t_pg = t0;
                        grad_prev = false;
                        continue;
                    }
                    if (err_new <= err + 2.0 * alpha * gd) {
                        found = true;
                        break;
                    }
                }
                if (stop) break;
                if (!found) {
                    grad_prev = false;
                    break;
                }
                grad_prev = true;
            }
            dp_sq = 0.0;
            for (int i = 0; i < kM; i++) dp_sq += (q[i] - p[i]) * (q[i] - p[i]);
            if (dp_sq <= eps2_sq * p_sq) {
                stop = true;
                break;
            }
            for (int i = 0; i < kM; i++) p[i] = q[i];
            memcpy(e, hx, sizeof(double) * n);
            err = err_new;
            break;
        }
    }
    delete[] buf;
}

}  // namespace wp
