// Homogeneous self-dual interior-point driver: a restatement of HELP's
// HSDSolver.cpp (the caller of the normal-equation backend) over an explicit
// help_lp and any help_backend.  Operation order, tolerances and status
// logic follow the reference line by line so that, with the same backend,
// the iterates are identical to the reference driver's.
//
// Column layout (Presolve.cpp:1101-1170): x[0, n_ub) have lower and upper
// bounds, x[n_ub, n_lb) a lower bound only, x[n_lb, n_lb + n_fr) are free.
//
// Two summation modes: the exact mode sums every reduction sequentially as
// the reference does (bit-identical iterates); the default mode splits long
// sums into 64 fixed chunks summed in parallel and then in chunk order, so
// its results are deterministic and independent of the thread count.  The
// ratio test's minimum is order independent in both.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "help_mi355x.h"

namespace {

bool g_exact = true;          // summation mode of the current help_hsd_solve call
constexpr int kChunks = 64;
constexpr int kParMin = 65536; // shorter sums stay sequential in both modes

template <class F>
double rsum(int n, F f) {
    if (g_exact || n < kParMin) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += f(i);
        return s;
    }
    double part[kChunks];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < kChunks; ++k) {
        int lo = (int)((int64_t)n * k / kChunks), hi = (int)((int64_t)n * (k + 1) / kChunks);
        double s = 0;
        for (int i = lo; i < hi; ++i) s += f(i);
        part[k] = s;
    }
    double s = 0;
    for (int k = 0; k < kChunks; ++k) s += part[k];
    return s;
}

// LP.h:81-96
constexpr double kStepGamma = 0.99995;
constexpr double kStepBeta = 1e-8;
constexpr double kMuTol = 1e-12;
constexpr double kInfeasTol = 1e-8;
constexpr double kGapTol = 1e-8;
constexpr double kPinfTol = 1e-8;
constexpr double kDinfTol = 1e-8;
constexpr double kVarTol = 1e-12;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double dot(int n, const double* a, const double* b) {   // Helper.cpp:48-54
    return rsum(n, [&](int i) { return a[i] * b[i]; });
}

struct Hsd {
    const help_lp* lp;
    const help_backend* be;
    int m, n, n_lb, n_ub, n_fr;
    const double *c, *b, *u;
    std::vector<double> x, y, z, xu, yu, zu, xf;
    double tau = 1, kappa = 1, rho = 1, inf0 = 1;
    double sum_r_norm = 0, pobj = 0, dobj = 0, comp_gap = 0, rel_gap = 0, pinf = 0, dinf = 0, infe = 0;
    int status = HELP_HSD_OK;
    int stall_reason = HELP_STALL_NONE;
    double newton_ratio = 0;
    std::vector<double> D, D_u, r_p, r_d;
    double D_g = 0, r_g = 0, mu = 0, rfval = 0;
    std::vector<double> sle1, sle2;                       // HSD_SLE_RHS1/2
    std::vector<double> bot1, bot1u, bot2, bot2u;         // HSD_LEBOT_*
    std::vector<double> dir1, dir1u, dir2, dir2u;         // HSD_Dir_RHS*
    std::vector<double> pp, ppu, qq, qqu;                 // HSD_Dir_p, p_u, q, q_u
    std::vector<double> uu, uuu, vv, vvu;                 // HSD_Dir_u, u_u, v, v_u
    std::vector<double> hxz, hxuzu;
    double htk = 0;
    std::vector<double> dx, dxu, dy, dyu, dz, dzu, dx_b, dxu_b, dy_b, dyu_b, dz_b, dzu_b;
    double dtau = 0, dkappa = 0;
    int backend_error = 0;
    double t_renew = 0, t_solve = 0, t_newton = 0, t_direction = 0, t_step = 0;
    struct Timer {   // adds the scope's wall time to *acc
        double* acc;
        double t0;
        explicit Timer(double* a) : acc(a), t0(now_s()) {}
        ~Timer() { *acc += now_s() - t0; }
    };
    int n_renew = 0, n_solve = 0;

    // A by rows (columns ascending), for the row-parallel forms of the
    // reference's column-order scatters: each row receives its terms in the
    // same order, so the results are bit-identical.
    std::vector<int64_t> rptr;
    std::vector<int32_t> rcol;
    std::vector<double> rval;

    explicit Hsd(const help_lp* l, const help_backend* b_) : lp(l), be(b_) {
        m = l->m; n = l->n; n_lb = l->n_lb; n_ub = l->n_ub; n_fr = l->n_fr;
        c = l->c; b = l->b; u = l->ub;
        const int64_t nnz = l->colptr[n];
        rptr.assign(m + 1, 0);
        for (int64_t p = 0; p < nnz; ++p) rptr[l->rowind[p] + 1]++;
        for (int i = 0; i < m; ++i) rptr[i + 1] += rptr[i];
        rcol.resize(nnz);
        rval.resize(nnz);
        std::vector<int64_t> fill(rptr.begin(), rptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int64_t p = l->colptr[j]; p < l->colptr[j + 1]; ++p) {
                int64_t t = fill[l->rowind[p]]++;
                rcol[t] = j;
                rval[t] = l->val[p];
            }
        auto z0 = [](std::vector<double>& v, int k) { v.assign(std::max(k, 1), 0.0); };
        z0(x, n); z0(y, m); z0(z, n); z0(xu, n_ub); z0(yu, n_ub); z0(zu, n_ub); z0(xf, n_fr);
        z0(D, n); z0(D_u, n_ub); z0(r_p, m); z0(r_d, n);
        z0(sle1, n); z0(sle2, m);
        z0(bot1, n); z0(bot1u, n_ub); z0(bot2, m); z0(bot2u, n_ub);
        z0(dir1, n); z0(dir1u, n_ub); z0(dir2, m); z0(dir2u, n_ub);
        z0(pp, n); z0(ppu, n_ub); z0(qq, m); z0(qqu, n_ub);
        z0(uu, n); z0(uuu, n_ub); z0(vv, m); z0(vvu, n_ub);
        z0(hxz, n); z0(hxuzu, n_ub);
        z0(dx, n); z0(dxu, n_ub); z0(dy, m); z0(dyu, n_ub); z0(dz, n); z0(dzu, n_ub);
        z0(dx_b, n); z0(dxu_b, n_ub); z0(dy_b, m); z0(dyu_b, n_ub); z0(dz_b, n); z0(dzu_b, n_ub);
    }

    // Helper.cpp:110-155 (SetATimesVector); dest_i receives its terms in
    // ascending column order, exactly as the reference's column scatter does
    void a_times(int sign, const double* v, double* dest) const {
#pragma omp parallel for schedule(static, 4096)
        for (int i = 0; i < m; ++i) {
            double acc = dest[i];
            for (int64_t t = rptr[i]; t < rptr[i + 1]; ++t) {
                if (sign == 1) acc += rval[t] * v[rcol[t]];
                else acc -= rval[t] * v[rcol[t]];
            }
            dest[i] = acc;
        }
    }
    void at_times(int sign, const double* v, double* dest) const {
#pragma omp parallel for schedule(static, 4096)
        for (int j = 0; j < n; ++j) {
            double dj = dest[j];
            for (int64_t p = lp->colptr[j]; p < lp->colptr[j + 1]; ++p) {
                if (sign == 1) dj += lp->val[p] * v[lp->rowind[p]];
                else dj -= lp->val[p] * v[lp->rowind[p]];
            }
            dest[j] = dj;
        }
    }
    static void scaled(int k, double alpha, const double* src, double* dst) {   // Helper.cpp:83-105
        if (alpha == 1) for (int i = 0; i < k; ++i) dst[i] = src[i];
        else if (alpha == -1) for (int i = 0; i < k; ++i) dst[i] = -src[i];
        else if (alpha == 0) for (int i = 0; i < k; ++i) dst[i] = 0;
        else for (int i = 0; i < k; ++i) dst[i] = alpha * src[i];
    }

    void renew() {
        double t0 = now_s();
        int st = be->renew(be->ctx, D.data());
        t_renew += now_s() - t0;
        ++n_renew;
        if (st && !backend_error) backend_error = st;
    }
    void solve(const double* b1, const double* b2, double* x1, double* x2) {
        double t0 = now_s();
        int st = be->solve(be->ctx, D.data(), b1, b2, x1, x2);
        t_solve += now_s() - t0;
        ++n_solve;
        if (st && !backend_error) backend_error = st;
    }

    // HSDSolver.cpp:53-150
    // min and sum of (a_i + sa*da_i) * (b_i + sb*db_i) over i < k, chunked
    void prod_min_sum(int k, const double* a, double sa, const double* da, const double* b, double sb,
                      const double* db, double& mn, double& sm) const {
        double pmin[kChunks], psum[kChunks];
#pragma omp parallel for schedule(static)
        for (int q = 0; q < kChunks; ++q) {
            int lo = (int)((int64_t)k * q / kChunks), hi = (int)((int64_t)k * (q + 1) / kChunks);
            double mloc = HUGE_VAL, sloc = 0;
            for (int i = lo; i < hi; ++i) {
                double t = da ? (a[i] + sa * da[i]) * (b[i] + sb * db[i]) : a[i] * b[i];
                mloc = std::min(mloc, t);
                sloc += t;
            }
            pmin[q] = mloc;
            psum[q] = sloc;
        }
        double s = 0;
        for (int q = 0; q < kChunks; ++q) { mn = std::min(mn, pmin[q]); s += psum[q]; }
        sm += s;
    }

    void newton_parameters() {
        Timer tm(&t_newton);
        double xz_min = tau * kappa, xz_sum = xz_min;
        if (!g_exact && n_lb >= kParMin) {
            bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
            for (int i = 0; i < n_lb; ++i) {
                D[i] = z[i] / x[i];
                if (D[i] <= 0.0) bad = true;
            }
            if (bad) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_NEGATIVE; return; }
            prod_min_sum(n_lb, z.data(), 0, nullptr, x.data(), 0, nullptr, xz_min, xz_sum);
        } else
        for (int i = 0; i < n_lb; ++i) {
            D[i] = z[i] / x[i];
            double t = z[i] * x[i];
            xz_min = std::min(xz_min, t);
            xz_sum += t;
            if (D[i] <= 0.0) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_NEGATIVE; return; }
        }
        for (int i = 0; i < n_ub; ++i) {
            D_u[i] = zu[i] / xu[i];
            D[i] += D_u[i];
            double t = zu[i] * xu[i];
            xz_min = std::min(xz_min, t);
            xz_sum += t;
            if (D_u[i] <= 0.0) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_NEGATIVE; return; }
        }
        double tmp = 0;
        for (int i = 0; i < n_fr; ++i) {
            D[n_lb + i] = rho * xf[i];
            tmp += xf[i] * x[n_lb + i] * x[n_lb + i];
        }
        D_g = kappa / tau + rho * tmp / (tau * tau);
        mu = xz_sum / (1.0 + n_lb + n_ub);
        rfval = (1.0 + n_lb + n_ub) * xz_min / xz_sum;
        scaled(m, tau, b, r_p.data());
        a_times(-1, x.data(), r_p.data());
        for (int i = 0; i < n_lb; ++i) r_d[i] = z[i] - c[i] * tau;
        for (int i = 0; i < n_ub; ++i) r_d[i] += yu[i];
        for (int i = 0; i < n_fr; ++i) r_d[n_lb + i] = -c[n_lb + i] * tau;
        for (int i = n_lb + n_fr; i < n; ++i) r_d[i] = 0;
        at_times(1, y.data(), r_d.data());
        double cTx = dot(n_lb + n_fr, c, x.data());
        double bTy = dot(m, b, y.data()) + dot(n_ub, u, yu.data());
        r_g = kappa + cTx - bTy;
        double rp2 = dot(m, r_p.data(), r_p.data());
        for (int i = 0; i < n_ub; ++i) rp2 += (tau * u[i] - xu[i] - x[i]) * (tau * u[i] - xu[i] - x[i]);
        double rd2 = dot(n_lb + n_fr, r_d.data(), r_d.data());
        for (int i = 0; i < n_ub; ++i) rd2 += (zu[i] + yu[i]) * (zu[i] + yu[i]);
        double rg2 = r_g * r_g;
        sum_r_norm = rp2 + rd2 + rg2;
        pobj = lp->c_intercept + cTx / tau;
        dobj = lp->c_intercept + bTy / tau;
        comp_gap = xz_sum;
        rel_gap = std::fabs(pobj - dobj) / (1.0 + std::fabs(dobj));
        pinf = std::sqrt(rp2) / (tau + std::sqrt(dot(n, x.data(), x.data())));
        dinf = std::sqrt(rd2) / (tau + std::sqrt(dot(n_lb, z.data(), z.data()) + dot(n_ub, zu.data(), zu.data()) +
                                                  dot(n_ub, yu.data(), yu.data()) + dot(m, y.data(), y.data())));
        infe = tau / (kappa * inf0);
    }

    // HSDSolver.cpp:156-208
    void solve_augmented(const double* s, const double* s_u, const double* t, const double* t_u,
                         double* p, double* p_u, double* q, double* q_u) {
        for (int i = 0; i < n_lb + n_fr; ++i) sle1[i] = s[i];
        for (int i = 0; i < n_ub; ++i) sle1[i] += D_u[i] * t_u[i] - s_u[i];
        for (int i = 0; i < m; ++i) sle2[i] = t[i];
        solve(sle1.data(), sle2.data(), p, q);
        for (int i = 0; i < m; ++i) q[i] = -q[i];
        for (int i = 0; i < n_ub; ++i) {
            p_u[i] = t_u[i] - p[i];
            q_u[i] = D_u[i] * p_u[i] - s_u[i];
        }
    }

    // HSDSolver.cpp:212-220
    double mix_dots(const double* p, const double* q, const double* q_u) const {
        double r = -dot(n_lb + n_fr, c, p);
        double rt = rho / tau;
        for (int i = 0; i < n_fr; ++i) r -= rt * xf[i] * x[i + n_lb] * p[i + n_lb];
        r += dot(m, b, q) + dot(n_ub, u, q_u);
        return r;
    }

    // HSDSolver.cpp:226-261
    double bottom() {
        for (int i = 0; i < n; ++i) bot1[i] = c[i];
        double rt = rho / tau;
        for (int i = 0; i < n_fr; ++i) bot1[n_lb + i] -= rt * xf[i] * x[n_lb + i];
        for (int i = 0; i < m; ++i) bot2[i] = -b[i];
        for (int i = 0; i < n_ub; ++i) { bot1u[i] = 0; bot2u[i] = -u[i]; }
        solve_augmented(bot1.data(), bot1u.data(), bot2.data(), bot2u.data(), pp.data(), ppu.data(),
                        qq.data(), qqu.data());
        return mix_dots(pp.data(), qq.data(), qqu.data());
    }

    // HSDSolver.cpp:281-473
    int direction(double eta, double DOT, double* hrxz, double* hrxuzu, double hrtk, double* odx,
                  double* odxu, double& odtau, double* ody, double* odyu, double* odz, double* odzu,
                  double& odkappa, double& rpn, double& rdn, double& rgn) {
        // wall time of the direction minus the backend solve inside it
        struct SolveSub {
            double* dir;
            double* solve;
            double s0;
            ~SolveSub() { *dir -= *solve - s0; }
        };
        Timer tm(&t_direction);
        SolveSub sub{&t_direction, &t_solve, t_solve};
        scaled(n, eta, r_d.data(), dir1.data());
        scaled(m, eta, r_p.data(), dir2.data());
        for (int i = 0; i < n_ub; ++i) {
            dir1u[i] = eta * (zu[i] + yu[i]);
            dir2u[i] = eta * (tau * u[i] - x[i] - xu[i]);
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n_lb; ++i) hrxz[i] /= std::max(x[i], z[i]);
        for (int i = 0; i < n_ub; ++i) hrxuzu[i] /= std::max(xu[i], zu[i]);
        for (int i = 0; i < n_lb; ++i)
            if (x[i] >= z[i]) dir1[i] += hrxz[i];
        for (int i = 0; i < n_ub; ++i)
            if (xu[i] >= zu[i]) dir1u[i] += hrxuzu[i];
        // HSDSolver.cpp:339-344 as a row gather (same per-row term order)
#pragma omp parallel for schedule(static, 4096)
        for (int r = 0; r < m; ++r) {
            double acc = dir2[r];
            for (int64_t t = rptr[r]; t < rptr[r + 1]; ++t) {
                int j = rcol[t];
                if (j < n_lb && x[j] < z[j]) acc -= rval[t] * hrxz[j];
            }
            dir2[r] = acc;
        }
        for (int i = 0; i < n_ub; ++i) {
            if (x[i] < z[i]) dir2u[i] -= hrxz[i];
            if (xu[i] < zu[i]) dir2u[i] -= hrxuzu[i];
        }
        double UP = eta * r_g + hrtk / tau;
        if (!g_exact && n_lb >= kParMin)
            UP += rsum(n_lb, [&](int i) { return x[i] < z[i] ? c[i] * hrxz[i] : 0.0; });
        else
        for (int i = 0; i < n_lb; ++i)
            if (x[i] < z[i]) UP += c[i] * hrxz[i];
        double DOWN = D_g - DOT;
        solve_augmented(dir1.data(), dir1u.data(), dir2.data(), dir2u.data(), uu.data(), uuu.data(),
                        vv.data(), vvu.data());
        if (DOWN <= 1e-13) odtau = 0;
        else odtau = (UP - mix_dots(uu.data(), vv.data(), vvu.data())) / std::max(D_g, DOWN);
        const double dt = odtau;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n_lb; ++i) {
            double T = uu[i] - pp[i] * dt;
            if (x[i] >= z[i]) { odx[i] = T; odz[i] = hrxz[i] - z[i] * T / x[i]; }
            else { odx[i] = hrxz[i] + T; odz[i] = -z[i] * T / x[i]; }
        }
        for (int i = 0; i < n_fr; ++i) {
            odx[i + n_lb] = uu[i + n_lb] - pp[i + n_lb] * dt;
            odz[i + n_lb] = hrxz[i + n_lb];
        }
        for (int i = 0; i < n_ub; ++i) {
            double T = uuu[i] - ppu[i] * dt;
            if (xu[i] >= zu[i]) { odxu[i] = T; odzu[i] = hrxuzu[i] - zu[i] * T / xu[i]; }
            else { odxu[i] = hrxuzu[i] + T; odzu[i] = -zu[i] * T / xu[i]; }
        }
        for (int i = 0; i < m; ++i) ody[i] = vv[i] - qq[i] * dt;
        for (int i = 0; i < n_ub; ++i) odyu[i] = vvu[i] - qqu[i] * dt;
        odkappa = (hrtk - kappa * dt) / tau;
        scaled(n, eta, r_d.data(), dir1.data());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n_lb; ++i) dir1[i] += -c[i] * dt + odz[i];
        double rt = rho / tau;
        for (int i = 0; i < n_fr; ++i)
            dir1[i + n_lb] += -(c[i] - rt * xf[i] * x[i + n_lb]) * dt - rho * xf[i] * odx[i + n_lb];
        at_times(1, ody, dir1.data());
        for (int i = 0; i < m; ++i) dir2[i] = eta * r_p[i] + b[i] * dt;
        a_times(-1, odx, dir2.data());
        for (int i = 0; i < n_ub; ++i) {
            dir1[i] += odyu[i];
            dir1u[i] = eta * (zu[i] + yu[i]) + odyu[i] + odzu[i];
            dir2u[i] = eta * (tau * u[i] - x[i] - xu[i]) + u[i] * dt - odxu[i] - odx[i];
        }
        rpn = dot(m, dir2.data(), dir2.data()) + dot(n_ub, dir2u.data(), dir2u.data());
        rdn = dot(n, dir1.data(), dir1.data()) + dot(n_ub, dir1u.data(), dir1u.data());
        double tmp = 0;
        for (int i = 0; i < n_fr; ++i) tmp += xf[i] * x[n_lb + i] * x[n_lb + i];
        rgn = eta * r_g + odkappa - mix_dots(odx, ody, odyu) - rho * tmp * dt / (tau * tau);
        rgn = rgn * rgn;
        if (rpn + rdn <= 1e-10 || rpn + rdn + rgn <= sum_r_norm * 0.1) return 1;
        return 0;
    }

    // HSDSolver.cpp:478-488
    double stepsize(const double* v, const double* vu, const double* dv, const double* dvu) const {
        double r = 1.0;
        if (!g_exact && n_lb >= kParMin) {
            // the sequential test keeps the smallest ratio -v/dv below 1
#pragma omp parallel for schedule(static) reduction(min : r)
            for (int i = 0; i < n_lb; ++i)
                if (dv[i] < -kVarTol && v[i] + dv[i] < 0.0) r = std::min(r, -v[i] / dv[i]);
        } else
        for (int i = 0; i < n_lb; ++i)
            if (dv[i] < -kVarTol && v[i] + r * dv[i] < 0.0) r = -v[i] / dv[i];
        for (int i = 0; i < n_ub; ++i)
            if (dvu[i] < -kVarTol && vu[i] + r * dvu[i] < 0.0) r = -vu[i] / dvu[i];
        return r;
    }

    // HSDSolver.cpp:491-514
    int check_step(double ap, double at, double ad, double ak, double beta) const {
        double xz_min = (tau + at * dtau) * (kappa + ak * dkappa);
        double xz_sum = xz_min;
        if (!g_exact && n_lb >= kParMin)
            prod_min_sum(n_lb, x.data(), ap, dx.data(), z.data(), ad, dz.data(), xz_min, xz_sum);
        else
        for (int i = 0; i < n_lb; ++i) {
            double t = (x[i] + ap * dx[i]) * (z[i] + ad * dz[i]);
            xz_min = std::min(xz_min, t);
            xz_sum += t;
        }
        for (int i = 0; i < n_ub; ++i) {
            double t = (xu[i] + ap * dxu[i]) * (zu[i] + ad * dzu[i]);
            xz_min = std::min(xz_min, t);
            xz_sum += t;
        }
        double mu_ = xz_sum / (1.0 + n_lb + n_ub);
        return xz_min / mu_ >= beta;
    }

    // HSDSolver.cpp:521-535
    void get_stepsize(double& ps, double& ds) {
        Timer tm(&t_step);
        ps = stepsize(x.data(), xu.data(), dx.data(), dxu.data());
        ds = stepsize(z.data(), zu.data(), dz.data(), dzu.data());
        if (dtau < -kVarTol) { ps = std::min(ps, -tau / dtau); ds = std::min(ds, -tau / dtau); }
        if (dkappa < -kVarTol) { ps = std::min(ps, -kappa / dkappa); ds = std::min(ds, -kappa / dkappa); }
    }

    // HSDSolver.cpp:537-594
    void update_step(double p_step, double d_step) {
        Timer tm(&t_step);
        double tau_step, kappa_step;
        int PGD = (p_step * dtau) > (d_step * dtau);
        if (PGD) { tau_step = d_step; kappa_step = p_step; }
        else { tau_step = p_step; kappa_step = d_step; }
        double rstep = kStepGamma;
        double threshold = std::max(std::min(std::sqrt(kStepBeta), rfval * 0.01), kStepBeta);
        while (!check_step(rstep * p_step, rstep * tau_step, rstep * d_step, rstep * kappa_step, threshold))
            rstep *= 0.95;
        p_step *= rstep; tau_step *= rstep; d_step *= rstep; kappa_step *= rstep;
        double tau_p = tau + p_step * dtau, tau_d = tau + d_step * dtau;
        double pf, df;
        if (PGD) { pf = tau_d / tau_p; df = 1; }
        else { pf = 1; df = tau_p / tau_d; }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n_lb; ++i) {
            x[i] = (x[i] + p_step * dx[i]) * pf;
            z[i] = (z[i] + d_step * dz[i]) * df;
        }
        for (int i = 0; i < m; ++i) y[i] = (y[i] + d_step * dy[i]) * df;
        for (int i = 0; i < n_ub; ++i) {
            xu[i] = (xu[i] + p_step * dxu[i]) * pf;
            yu[i] = (yu[i] + d_step * dyu[i]) * df;
            zu[i] = (zu[i] + d_step * dzu[i]) * df;
        }
        for (int i = 0; i < n_fr; ++i) x[i + n_lb] = (x[i + n_lb] + p_step * dx[i + n_lb]) * pf;
        tau += tau_step * dtau;
        kappa = (kappa + kappa_step * dkappa) * pf * df;
    }

    void take_improved(double dtau_b, double dkappa_b) {
        std::copy(dx_b.begin(), dx_b.begin() + n, dx.begin());
        std::copy(dxu_b.begin(), dxu_b.begin() + n_ub, dxu.begin());
        std::copy(dy_b.begin(), dy_b.begin() + m, dy.begin());
        std::copy(dyu_b.begin(), dyu_b.begin() + n_ub, dyu.begin());
        std::copy(dz_b.begin(), dz_b.begin() + n, dz.begin());
        std::copy(dzu_b.begin(), dzu_b.begin() + n_ub, dzu.begin());
        dtau = dtau_b;
        dkappa = dkappa_b;
    }

    // HSDSolver.cpp:596-717
    void init_point() {
        status = HELP_HSD_OK;
        for (int i = 0; i < n_lb; ++i) x[i] = z[i] = 1.0;
        for (int i = 0; i < n_ub; ++i) { xu[i] = zu[i] = 1.0; yu[i] = 0.0; }
        for (int i = 0; i < m; ++i) y[i] = 0.0;
        tau = kappa = 1.0;
        for (int i = 0; i < n_fr; ++i) {
            x[n_lb + i] = z[n_lb + i] = 0.0;
            xf[i] = 1.0 / (1.0 + 10.0 * std::fabs(x[n_lb + i]));
        }
        double nall = 1.0 + n_lb + n_ub;
        double mu0 = (dot(n_lb, x.data(), z.data()) + dot(n_ub, xu.data(), zu.data()) + tau * kappa) / nall;
        inf0 = tau / kappa;
        rho = 1.0;
        newton_parameters();
        if (status != HELP_HSD_OK) return;
        renew();
        double DOT = bottom();
        double rpn, rdn, rgn;
        double phi = 1.0, eta = 1.0;
        for (int i = 0; i < n_lb; ++i) hxz[i] = -x[i] * z[i] + mu0;
        for (int i = 0; i < n_fr; ++i) hxz[i + n_lb] = 0;
        for (int i = 0; i < n_ub; ++i) hxuzu[i] = -xu[i] * zu[i] + mu0;
        htk = -tau * kappa + mu0;
        int newton_ok = direction(eta, DOT, hxz.data(), hxuzu.data(), htk, dx.data(), dxu.data(), dtau,
                                  dy.data(), dyu.data(), dz.data(), dzu.data(), dkappa, rpn, rdn, rgn);
        double p_step = 0, d_step = 0;
        if (newton_ok) get_stepsize(p_step, d_step);
        else phi = 0.0;
        double gamma = 10.0;
        double alpha_p = std::max(0.0, 1.0 - std::min(p_step, d_step)) * gamma * mu0;
        double dmu = std::max((dot(n_lb, dx.data(), dz.data()) + dot(n_ub, dxu.data(), dzu.data()) + dtau * dkappa) / nall, 0.0);
        for (int i = 0; i < n_lb; ++i) hxz[i] = -x[i] * z[i] + alpha_p - phi * (dx[i] * dz[i] - dmu);
        for (int i = 0; i < n_fr; ++i) hxz[i + n_lb] = 0;
        for (int i = 0; i < n_ub; ++i) hxuzu[i] = -xu[i] * zu[i] + alpha_p - phi * (dxu[i] * dzu[i] - dmu);
        htk = -tau * kappa + alpha_p - phi * (dtau * dkappa - dmu);
        double dtau_b, dkappa_b, rpb, rdb, rgb;
        int imp_ok = direction(eta, DOT, hxz.data(), hxuzu.data(), htk, dx_b.data(), dxu_b.data(), dtau_b,
                               dy_b.data(), dyu_b.data(), dz_b.data(), dzu_b.data(), dkappa_b, rpb, rdb, rgb);
        if (!imp_ok) {
            if (!newton_ok) return;
        } else {
            take_improved(dtau_b, dkappa_b);
            get_stepsize(p_step, d_step);
        }
        if (p_step < kVarTol && d_step < kVarTol) return;
        update_step(p_step, d_step);
    }

    // HSDSolver.cpp:737-867
    int solve_all(int max_iter, help_hsd_log_fn logf, void* user) {
        init_point();
        double mu0 = (dot(n_lb, x.data(), z.data()) + dot(n_ub, xu.data(), zu.data()) + tau * kappa) / (1.0 + n_lb + n_ub);
        mu = mu0;
        int it;
        for (it = 0; it < max_iter; ++it) {
            // a failed backend call (e.g. PCG stopped at its cap short of
            // its tolerance) ends the run: the reference would exit in
            // CheckError (Helper.cpp:24-35)
            if (backend_error) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_BACKEND; break; }
            rho = std::max(mu / mu0, 1e-12);
            for (int i = 0; i < n_fr; ++i) xf[i] = 1.0 / (1.0 + 10.0 * std::fabs(x[n_lb + i]));
            newton_parameters();
            if (logf) logf(user, it, pobj, dobj, rel_gap, pinf, dinf);
            if (status != HELP_HSD_OK || (mu <= kMuTol * mu0 && infe < kInfeasTol) ||
                (rel_gap <= kGapTol && pinf <= kPinfTol && dinf <= kDinfTol))
                break;
            renew();
            double DOT = bottom();
            double rpn, rdn, rgn;
            double phi = 1.0, eta = 1.0;
            for (int i = 0; i < n_lb; ++i) hxz[i] = -x[i] * z[i];
            for (int i = 0; i < n_fr; ++i) hxz[i + n_lb] = 0;
            for (int i = 0; i < n_ub; ++i) hxuzu[i] = -xu[i] * zu[i];
            htk = -tau * kappa;
            int newton_ok = direction(eta, DOT, hxz.data(), hxuzu.data(), htk, dx.data(), dxu.data(), dtau,
                                      dy.data(), dyu.data(), dz.data(), dzu.data(), dkappa, rpn, rdn, rgn);
            double p_step = 0, d_step = 0, gamma;
            if (newton_ok) {
                get_stepsize(p_step, d_step);
                double new_mu = (tau + p_step * dtau) * (kappa + d_step * dkappa);
                if (!g_exact && n_lb >= kParMin) {
                    double mn = HUGE_VAL;
                    prod_min_sum(n_lb, x.data(), p_step, dx.data(), z.data(), d_step, dz.data(), mn, new_mu);
                } else
                for (int i = 0; i < n_lb; ++i) new_mu += (x[i] + p_step * dx[i]) * (z[i] + d_step * dz[i]);
                for (int i = 0; i < n_ub; ++i) new_mu += (xu[i] + p_step * dxu[i]) * (zu[i] + d_step * dzu[i]);
                new_mu /= 1.0 + n_lb + n_ub;
                gamma = new_mu / mu;
                gamma = std::min(gamma * gamma * gamma, 0.2 * std::min(gamma, 1.0));
                if (mu < 1e-3) gamma = std::max(gamma, mu);
                gamma = std::max(1e-6, gamma);
            } else {
                status = HELP_HSD_STALLED;
                stall_reason = HELP_STALL_DIRECTION;
                newton_ratio = (rpn + rdn + rgn) / sum_r_norm;
                break;
            }
            eta = 1.0 - gamma;
            double alpha_p = gamma * mu;
            double dmu = std::max((dot(n_lb, dx.data(), dz.data()) + dot(n_ub, dxu.data(), dzu.data()) + dtau * dkappa) / (1.0 + n_lb + n_ub), 0.0);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n_lb; ++i) hxz[i] = -x[i] * z[i] + alpha_p - phi * (dx[i] * dz[i] - dmu);
            for (int i = 0; i < n_fr; ++i) hxz[i + n_lb] = 0;
            for (int i = 0; i < n_ub; ++i) hxuzu[i] = -xu[i] * zu[i] + alpha_p - phi * (dxu[i] * dzu[i] - dmu);
            htk = -tau * kappa + alpha_p - phi * (dtau * dkappa - dmu);
            double dtau_b, dkappa_b, rpb, rdb, rgb;
            int imp_ok = direction(eta, DOT, hxz.data(), hxuzu.data(), htk, dx_b.data(), dxu_b.data(), dtau_b,
                                   dy_b.data(), dyu_b.data(), dz_b.data(), dzu_b.data(), dkappa_b, rpb, rdb, rgb);
            if (imp_ok) {
                take_improved(dtau_b, dkappa_b);
                get_stepsize(p_step, d_step);
            }
            if (p_step < kVarTol && d_step < kVarTol) {
                status = HELP_HSD_UNKNOWN_ERROR;
                break;
            }
            update_step(p_step, d_step);
        }
        if (it >= max_iter) status = HELP_HSD_ITER_LIMIT;
        return it;
    }
};

}  // namespace

extern "C" int help_hsd_solve(const help_lp* lp, const help_backend* be, int32_t max_iter,
                              int32_t flags, help_hsd_log_fn logf, void* user, help_hsd_result* res) {
    if (!lp || !be || !res || !be->construct || !be->renew || !be->solve || !be->destruct)
        return HELP_ERR_ARG;
    g_exact = (flags & HELP_HSD_EXACT) != 0;
    if (lp->n_lb + lp->n_fr != lp->n || lp->n_ub > lp->n_lb || (lp->n_ub > 0 && !lp->ub))
        return HELP_ERR_ARG;
    std::memset(res, 0, sizeof(*res));
    if (max_iter <= 0) max_iter = 10000;   // LP.h:84 Max_Iterations
    double t0 = now_s();
    int st = be->construct(be->ctx, lp);
    res->seconds_construct = now_s() - t0;
    if (st) {
        res->backend_error = st;
        res->seconds = now_s() - t0;
        return st;
    }
    Hsd h(lp, be);
    int it = h.solve_all(max_iter, logf, user);
    be->destruct(be->ctx);
    res->seconds = now_s() - t0;
    res->status = h.status;
    res->stall_reason = h.stall_reason;
    res->newton_ratio = h.newton_ratio;
    res->iterations = it;
    res->primal_obj = h.pobj;
    res->dual_obj = h.dobj;
    res->rel_gap = h.rel_gap;
    res->primal_infeas = h.pinf;
    res->dual_infeas = h.dinf;
    res->mu = h.mu;
    res->seconds_renew = h.t_renew;
    res->seconds_solve = h.t_solve;
    res->seconds_newton = h.t_newton;
    res->seconds_direction = h.t_direction;
    res->seconds_step = h.t_step;
    res->n_renew = h.n_renew;
    res->n_solve = h.n_solve;
    res->backend_error = h.backend_error;
    return HELP_OK;
}
