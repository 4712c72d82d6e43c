// Device-resident homogeneous self-dual driver (restatement of HELP's
// HSDSolver.cpp, as hsd.cpp, with every vector on the GPU).
//
// The host keeps only the scalar logic (tau, kappa, mu, step lengths, the
// stop test); every O(n) loop, every product with A or A^T and every
// reduction runs on the MI355X in FP64: the D = Z/X update, the residuals
// r_p, r_d, the search directions, the ratio test and the step update.
// Reductions are deterministic two-stage sums (a fixed grid of block
// partials, then one pass in block order), so a run is reproducible; their
// order differs from the reference's sequential sums, so iterates match the
// host driver to rounding, not bitwise.  The backend is called with device
// pointers, so no vector crosses PCIe during the iterations.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "handle.h"

namespace help {

// acc + sum_p A_pj v[row_p] over column j in ascending p, with the loads of
// four terms in flight (indices past the column clamped, terms dropped)
__device__ inline double col_dot_acc(const int64_t* cptr, const int32_t* crow, const double* cval,
                                     const double* v, int64_t j, double acc) {
    const int64_t b = cptr[j], e = cptr[j + 1];
    for (int64_t p = b; p < e; p += 4) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t q = p + u < e ? p + u : e - 1;
            t[u] = cval[q] * v[crow[q]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = (p + u < e) ? acc + t[u] : acc;
    }
    return acc;
}

namespace {

constexpr int kRedBlocks = 512;

// LP.h:81-96
constexpr double kStepGamma = 0.99995;
constexpr double kStepBeta = 1e-8;
constexpr double kMuTol = 1e-12;
constexpr double kInfeasTol = 1e-8;
constexpr double kGapTol = 1e-8;
constexpr double kPinfTol = 1e-8;
constexpr double kDinfTol = 1e-8;
constexpr double kVarTol = 1e-12;

template <class F>
__global__ __launch_bounds__(256) void k_map(int64_t n, F f) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) f(i);
}

// K sums of f's outputs and the minimum of its first output
template <int K, class F>
__global__ __launch_bounds__(256) void k_reduce(int64_t n, F f, double* __restrict__ part) {
    __shared__ double red[256];
    double s[K], mn = HUGE_VAL;
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v[K];
        f(i, v);
        for (int k = 0; k < K; ++k) s[k] = s[k] + v[k];
        mn = fmin(mn, v[0]);
    }
    for (int k = 0; k <= K; ++k) {
        red[threadIdx.x] = k < K ? s[k] : mn;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) {
                double a = red[threadIdx.x], b = red[threadIdx.x + off];
                red[threadIdx.x] = k < K ? a + b : fmin(a, b);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (K + 1) + k] = red[0];
        __syncthreads();
    }
}

// Combines the per-block partials of k_reduce: thread t folds partials
// t, t+256, ... and a fixed LDS tree finishes, so the order depends only on
// the partial count.
template <int K>
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
    __shared__ double sm[256];
    const int t = threadIdx.x;
    for (int k = 0; k <= K; ++k) {
        const bool is_min = (k == K);
        double acc = is_min ? HUGE_VAL : 0.0;
        for (int b = t; b < nb; b += 256) {
            const double v = part[(int64_t)b * (K + 1) + k];
            acc = is_min ? fmin(acc, v) : acc + v;
        }
        sm[t] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) sm[t] = is_min ? fmin(sm[t], sm[t + w]) : sm[t] + sm[t + w];
            __syncthreads();
        }
        if (t == 0) out[k] = sm[0];
        __syncthreads();
    }
}

struct HsdGpu {
    const help_lp* lp;
    const help_backend* be;
    help_gpu* h = nullptr;
    hipStream_t st = nullptr;
    int m, n, n_lb, n_ub, n_fr;
    // device: problem data
    const int64_t *rowptr, *colptr;
    const int32_t *colind, *rowind;
    const double *csrval, *cscval;
    DBuf<double> c_, b_, u_;
    // device: iterates and work vectors
    DBuf<double> x, y, z, xu, yu, zu, xf, D, D_u, r_p, r_d, sle1, sle2, sle1b, sle2b, bot1, bot1u, bot2, bot2u, dir1,
        dir1u, dir2, dir2u, pp, ppu, qq, qqu, uu, uuu, vv, vvu, hxz, hxuzu, dx, dxu, dy, dyu, dz, dzu,
        dx_b, dxu_b, dy_b, dyu_b, dz_b, dzu_b, part, out, tmpm, tmpn;
    double tau = 1, kappa = 1, rho = 1, inf0 = 1;
    double sum_r_norm = 0, pobj = 0, dobj = 0, rel_gap = 0, pinf = 0, dinf = 0, infe = 0;
    int status = HELP_HSD_OK;
    int stall_reason = HELP_STALL_NONE;
    double newton_ratio = 0;
    double D_g = 0, r_g = 0, mu = 0, rfval = 0, htk = 0, dtau = 0, dkappa = 0;
    int backend_error = 0;
    double t_renew = 0, t_solve = 0, t_newton = 0, t_direction = 0, t_step = 0;
    int n_renew = 0, n_solve = 0;
    static constexpr int kOutSlots = 128;
    double* hout = nullptr;   // pinned host copy of `out`
    int used = 0;             // pending slots of `out`
    std::vector<std::pair<int, double>> empty_slots;

    struct Timer {
        double* acc;
        double t0;
        explicit Timer(double* a) : acc(a), t0(now_s()) {}
        ~Timer() { *acc += now_s() - t0; }
    };

    template <class F>
    void map(int64_t k, F f) {
        if (k > 0) help::launch_kernel(k_map<F>, (unsigned)((k + 255) / 256), 256, 0, st, k, f);
    }
    // Deferred reductions: reduce_late launches the reduction into its own
    // slot of `out` and returns the slot; flush() brings every pending slot
    // to pinned host memory with one copy and one synchronization, after
    // which val(slot, k) reads it (k < K: sum of output k; k = K: minimum of
    // output 0).  Independent reductions thus share one host round trip.
    template <int K, class F>
    int reduce_late(int64_t k, F f) {
        const int slot = used;
        if (used + K + 1 > kOutSlots) throw std::runtime_error("hsd_gpu: too many pending reductions");
        used += K + 1;
        if (k <= 0) {   // identities, written into the host copy by flush()
            for (int i = 0; i <= K; ++i) empty_slots.push_back({slot + i, i < K ? 0.0 : HUGE_VAL});
            return slot;
        }
        int nb = (int)std::min<int64_t>(kRedBlocks, (k + 255) / 256);
        help::launch_kernel(k_reduce<K, F>, nb, 256, 0, st, k, f, part.p);
        help::launch_kernel(k_reduce_final<K>, 1, 256, 0, st, part.p, nb, out.p + slot);
        return slot;
    }
    int dot_late(int64_t k, const double* a, const double* b) {
        return reduce_late<1>(k, [=] __device__(int64_t i, double* v) { v[0] = a[i] * b[i]; });
    }
    void flush() {
        if (used > 0) {
            HIPCHK(hipMemcpyAsync(hout, out.p, sizeof(double) * used, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        for (const auto& e : empty_slots) hout[e.first] = e.second;
        empty_slots.clear();
        used = 0;
    }
    double val(int slot, int k = 0) const { return hout[slot + k]; }
    double dot(int64_t k, const double* a, const double* b) {
        const int slot = dot_late(k, a, b);
        flush();
        return val(slot);
    }

    explicit HsdGpu(const help_lp* l, const help_backend* b, help_gpu* hh) : lp(l), be(b), h(hh) {
        st = h->st;
        m = l->m; n = l->n; n_lb = l->n_lb; n_ub = l->n_ub; n_fr = l->n_fr;
        rowptr = h->d_rowptr.p; colind = h->d_colind.p; csrval = h->d_csrval.p;
        colptr = h->d_colptr.p; rowind = h->d_rowind.p; cscval = h->d_val.p;
        auto up = [&](DBuf<double>& d, const double* src, int k) {
            d.alloc(std::max(k, 1));
            if (k > 0 && src) HIPCHK(hipMemcpyAsync(d.p, src, sizeof(double) * k, hipMemcpyHostToDevice, st));
        };
        up(c_, l->c, n); up(b_, l->b, m); up(u_, l->ub, n_ub);
        for (DBuf<double>* v : {&x, &z, &D, &r_d, &sle1, &sle1b, &bot1, &dir1, &pp, &uu, &hxz, &dx, &dz, &dx_b, &dz_b})
            v->alloc(std::max(n, 1));
        for (DBuf<double>* v : {&y, &r_p, &sle2, &sle2b, &bot2, &dir2, &qq, &vv, &dy, &dy_b}) v->alloc(std::max(m, 1));
        for (DBuf<double>* v : {&xu, &yu, &zu, &D_u, &bot1u, &bot2u, &dir1u, &dir2u, &ppu, &qqu, &uuu, &vvu,
                                &hxuzu, &dxu, &dyu, &dzu, &dxu_b, &dyu_b, &dzu_b})
            v->alloc(std::max(n_ub, 1));
        xf.alloc(std::max(n_fr, 1));
        tmpm.alloc(std::max(m, 1));
        tmpn.alloc(std::max(n, 1));
        part.alloc((size_t)kRedBlocks * 8);
        out.alloc(kOutSlots);
        HIPCHK(hipHostMalloc((void**)&hout, sizeof(double) * kOutSlots, hipHostMallocDefault));
        HIPCHK(hipStreamSynchronize(st));
    }
    ~HsdGpu() {
        if (hout) (void)hipHostFree(hout);
    }
    HsdGpu(const HsdGpu&) = delete;
    HsdGpu& operator=(const HsdGpu&) = delete;

    void renew() {
        Timer t(&t_renew);
        ++n_renew;
        int s = be->renew(be->ctx, D.p);
        if (s && !backend_error) backend_error = s;
    }
    void solve(const double* b1, const double* b2, double* x1, double* x2) {
        Timer t(&t_solve);
        ++n_solve;
        int s = be->solve(be->ctx, D.p, b1, b2, x1, x2);
        if (s && !backend_error) backend_error = s;
    }
    // two solves with the same D in one backend call (help_gpu_solve2)
    void solve2(const double* b1a, const double* b2a, double* x1a, double* x2a, const double* b1b,
                const double* b2b, double* x1b, double* x2b) {
        Timer t(&t_solve);
        n_solve += 2;
        int s = help_gpu_backend_solve2(const_cast<help_backend*>(be), D.p, b1a, b2a, x1a, x2a, b1b, b2b, x1b,
                                        x2b);
        if (s && !backend_error) backend_error = s;
    }

    // HSDSolver.cpp:53-150.  Every reduction is launched first and read
    // after one flush; the D updates and residual products that follow a
    // failed positivity check in the reference are harmless here, since the
    // driver stops on that status without reading D, r_p or r_d.
    void newton_parameters() {
        Timer tm(&t_newton);
        double* Dp = D.p; double* Dup = D_u.p;
        const double *xp = x.p, *zp = z.p, *xup = xu.p, *zup = zu.p, *yup = yu.p, *xfp = xf.p, *yp = y.p;
        const double *cp = c_.p, *bp = b_.p, *upp = u_.p;
        const double tau_ = tau, rho_ = rho;
        const int nlb = n_lb, nub = n_ub, nfr = n_fr;
        map(n_lb, [=] __device__(int64_t i) { Dp[i] = zp[i] / xp[i]; });
        const int o_r1 = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* v) { v[0] = zp[i] * xp[i]; });
        const int o_rd = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* v) { v[0] = Dp[i]; });
        int o_ru = -1, o_rdu = -1, o_tmp = -1, o_rp2u = -1, o_rd2u = -1;
        if (n_ub > 0) {
            map(n_ub, [=] __device__(int64_t i) { Dup[i] = zup[i] / xup[i]; Dp[i] += Dup[i]; });
            o_ru = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* v) { v[0] = zup[i] * xup[i]; });
            o_rdu = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* v) { v[0] = Dup[i]; });
        }
        if (n_fr > 0) {
            map(n_fr, [=] __device__(int64_t i) { Dp[nlb + i] = rho_ * xfp[i]; });
            o_tmp = reduce_late<1>(n_fr, [=] __device__(int64_t i, double* v) {
                v[0] = xfp[i] * xp[nlb + i] * xp[nlb + i];
            });
        }
        // r_p = tau b - A x  (A x by the row-split product: heavy rows in chunks)
        double* rpp = r_p.p;
        const double* ax = tmpm.p;
        launch_csr_split(h->rows, rowptr, colind, csrval, xp, tmpm.p, false, st);
        map(m, [=] __device__(int64_t i) { rpp[i] = tau_ * bp[i] - ax[i]; });
        // r_d = z - c tau (+yu) (free: -c tau), then += A^T y
        double* rdp = r_d.p;
        const int64_t* cptr = colptr; const int32_t* crow = rowind; const double* cval = cscval;
        map(n, [=] __device__(int64_t j) {
            double acc;
            if (j < nlb) acc = zp[j] - cp[j] * tau_;
            else if (j < nlb + nfr) acc = -cp[j] * tau_;
            else acc = 0.0;
            if (j < nub) acc += yup[j];
            rdp[j] = col_dot_acc(cptr, crow, cval, yp, j, acc);
        });
        const int o_ctx = dot_late(n_lb + n_fr, cp, xp);
        const int o_by = dot_late(m, bp, yp), o_uyu = dot_late(n_ub, upp, yup);
        const int o_rp2 = dot_late(m, rpp, rpp);
        if (n_ub > 0)
            o_rp2u = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* v) {
                double t = tau_ * upp[i] - xup[i] - xp[i];
                v[0] = t * t;
            });
        const int o_rd2 = dot_late(n_lb + n_fr, rdp, rdp);
        if (n_ub > 0)
            o_rd2u = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* v) {
                double t = zup[i] + yup[i];
                v[0] = t * t;
            });
        const int o_xx = dot_late(n, xp, xp);
        const int o_zz = dot_late(n_lb, zp, zp), o_zuzu = dot_late(n_ub, zup, zup);
        const int o_yuyu = dot_late(n_ub, yup, yup), o_yy = dot_late(m, yp, yp);
        flush();
        double xz_min = tau * kappa, xz_sum = xz_min;
        if (n_lb > 0 && val(o_rd, 1) <= 0.0) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_NEGATIVE; return; }
        xz_min = std::min(xz_min, val(o_r1, 1));
        xz_sum += val(o_r1);
        if (n_ub > 0) {
            if (val(o_rdu, 1) <= 0.0) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_NEGATIVE; return; }
            xz_min = std::min(xz_min, val(o_ru, 1));
            xz_sum += val(o_ru);
        }
        double tmp = 0;
        if (n_fr > 0) tmp = val(o_tmp);
        D_g = kappa / tau + rho * tmp / (tau * tau);
        mu = xz_sum / (1.0 + n_lb + n_ub);
        rfval = (1.0 + n_lb + n_ub) * xz_min / xz_sum;
        double cTx = val(o_ctx);
        double bTy = val(o_by) + val(o_uyu);
        r_g = kappa + cTx - bTy;
        double rp2 = val(o_rp2);
        if (n_ub > 0) rp2 += val(o_rp2u);
        double rd2 = val(o_rd2);
        if (n_ub > 0) rd2 += val(o_rd2u);
        double rg2 = r_g * r_g;
        sum_r_norm = rp2 + rd2 + rg2;
        pobj = lp->c_intercept + cTx / tau;
        dobj = lp->c_intercept + bTy / tau;
        rel_gap = std::fabs(pobj - dobj) / (1.0 + std::fabs(dobj));
        pinf = std::sqrt(rp2) / (tau + std::sqrt(val(o_xx)));
        dinf = std::sqrt(rd2) / (tau + std::sqrt(val(o_zz) + val(o_zuzu) + val(o_yuyu) + val(o_yy)));
        infe = tau / (kappa * inf0);
    }

    // HSDSolver.cpp:156-208 in two halves around the solve, so that two
    // systems can share one backend call: the right-hand side (s1, s2) ...
    void augmented_rhs(const double* s, const double* s_u, const double* t, const double* t_u, double* s1,
                       double* s2) {
        const double* Dup = D_u.p;
        const int nlbfr = n_lb + n_fr;
        map(nlbfr, [=] __device__(int64_t i) { s1[i] = s[i]; });
        map(n_ub, [=] __device__(int64_t i) { s1[i] += Dup[i] * t_u[i] - s_u[i]; });
        HIPCHK(hipMemcpyAsync(s2, t, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    }
    // ... and, with the solution (p, q) of that system, the remaining parts
    void augmented_finish(const double* s_u, const double* t_u, double* p, double* p_u, double* q,
                          double* q_u) {
        const double* Dup = D_u.p;
        map(m, [=] __device__(int64_t i) { q[i] = -q[i]; });
        map(n_ub, [=] __device__(int64_t i) {
            p_u[i] = t_u[i] - p[i];
            q_u[i] = Dup[i] * p_u[i] - s_u[i];
        });
    }

    // HSDSolver.cpp:212-220: mix_late launches the reductions, mix_value
    // combines them after a flush
    struct Mix { int cp, fr, bq, uq; };
    Mix mix_late(const double* p, const double* q, const double* q_u) {
        Mix r;
        r.cp = dot_late(n_lb + n_fr, c_.p, p);
        r.fr = -1;
        if (n_fr > 0) {
            const double* xfp = xf.p; const double* xp = x.p;
            const int nlb = n_lb;
            r.fr = reduce_late<1>(n_fr, [=] __device__(int64_t i, double* v) {
                v[0] = xfp[i] * xp[i + nlb] * p[i + nlb];
            });
        }
        r.bq = dot_late(m, b_.p, q);
        r.uq = dot_late(n_ub, u_.p, q_u);
        return r;
    }
    double mix_value(const Mix& x) const {
        double r = -val(x.cp);
        if (x.fr >= 0) r -= (rho / tau) * val(x.fr);
        r += val(x.bq) + val(x.uq);
        return r;
    }

    // HSDSolver.cpp:226-261: the right-hand side of the bottom system into
    // (sle1b, sle2b); bottom_finish() takes its solution from (pp, qq)
    void bottom_rhs() {
        double *b1 = bot1.p, *b1u = bot1u.p, *b2 = bot2.p, *b2u = bot2u.p;
        const double *cp = c_.p, *bp = b_.p, *upp = u_.p, *xfp = xf.p, *xp = x.p;
        const double rt = rho / tau;
        const int nlb = n_lb;
        map(n, [=] __device__(int64_t i) { b1[i] = cp[i]; });
        map(n_fr, [=] __device__(int64_t i) { b1[nlb + i] -= rt * xfp[i] * xp[nlb + i]; });
        map(m, [=] __device__(int64_t i) { b2[i] = -bp[i]; });
        map(n_ub, [=] __device__(int64_t i) { b1u[i] = 0; b2u[i] = -upp[i]; });
        augmented_rhs(b1, b1u, b2, b2u, sle1b.p, sle2b.p);
    }
    Mix bottom_finish() {
        augmented_finish(bot1u.p, bot2u.p, pp.p, ppu.p, qq.p, qqu.p);
        return mix_late(pp.p, qq.p, qqu.p);
    }

    // The bottom system and the first (affine) direction share D, and the
    // direction's right-hand side does not depend on the bottom solution,
    // so both systems go through one solve2 call; the arithmetic is the
    // same as solving them one after the other (HSDSolver.cpp:797-800).
    int bottom_and_direction(double eta, double& DOT, double* hrxz, double* hrxuzu, double hrtk, double* odx,
                             double* odxu, double& odtau, double* ody, double* odyu, double* odz, double* odzu,
                             double& odkappa, double& rpn, double& rdn, double& rgn) {
        const int o_up = direction_rhs(eta, hrxz, hrxuzu);
        solve2(sle1b.p, sle2b.p, pp.p, qq.p, sle1.p, sle2.p, uu.p, vv.p);
        const Mix bot = bottom_finish();
        return direction_finish(eta, &bot, DOT, o_up, hrxz, hrxuzu, hrtk, odx, odxu, odtau, ody, odyu, odz, odzu,
                                odkappa, rpn, rdn, rgn);
    }

    // HSDSolver.cpp:281-473
    int direction(double eta, double DOT, double* hrxz, double* hrxuzu, double hrtk, double* odx, double* odxu,
                  double& odtau, double* ody, double* odyu, double* odz, double* odzu, double& odkappa,
                  double& rpn, double& rdn, double& rgn) {
        const int o_up = direction_rhs(eta, hrxz, hrxuzu);
        solve(sle1.p, sle2.p, uu.p, vv.p);
        return direction_finish(eta, nullptr, DOT, o_up, hrxz, hrxuzu, hrtk, odx, odxu, odtau, ody, odyu, odz,
                                odzu, odkappa, rpn, rdn, rgn);
    }
    // the direction's right-hand side into (sle1, sle2); returns the slot of
    // the reduction that completes UP
    int direction_rhs(double eta, double* hrxz, double* hrxuzu) {
        Timer tm(&t_direction);
        double *d1 = dir1.p, *d1u = dir1u.p, *d2 = dir2.p, *d2u = dir2u.p;
        const double *rdp = r_d.p, *rpp = r_p.p, *xp = x.p, *zp = z.p, *xup = xu.p, *zup = zu.p,
                     *yup = yu.p, *cp = c_.p, *upp = u_.p;
        const double tau_ = tau;
        const int nlb = n_lb;
        map(n, [=] __device__(int64_t i) { d1[i] = eta * rdp[i]; });
        map(n_ub, [=] __device__(int64_t i) {
            d1u[i] = eta * (zup[i] + yup[i]);
            d2u[i] = eta * (tau_ * upp[i] - xp[i] - xup[i]);
        });
        map(n_lb, [=] __device__(int64_t i) {
            hrxz[i] /= fmax(xp[i], zp[i]);
            if (xp[i] >= zp[i]) d1[i] += hrxz[i];
        });
        map(n_ub, [=] __device__(int64_t i) {
            hrxuzu[i] /= fmax(xup[i], zup[i]);
            if (xup[i] >= zup[i]) d1u[i] += hrxuzu[i];
        });
        // dir2 = eta r_p - sum_{j: x_j < z_j} A(:,j) hrxz_j
        {
            double* hv = tmpn.p;
            const double* ah = tmpm.p;
            map(n, [=] __device__(int64_t j) { hv[j] = (j < nlb && xp[j] < zp[j]) ? hrxz[j] : 0.0; });
            launch_csr_split(h->rows, rowptr, colind, csrval, hv, tmpm.p, false, st);
            map(m, [=] __device__(int64_t r) { d2[r] = eta * rpp[r] - ah[r]; });
        }
        map(n_ub, [=] __device__(int64_t i) {
            if (xp[i] < zp[i]) d2u[i] -= hrxz[i];
            if (xup[i] < zup[i]) d2u[i] -= hrxuzu[i];
        });
        const int o_up = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* v) {
            v[0] = xp[i] < zp[i] ? cp[i] * hrxz[i] : 0.0;
        });
        augmented_rhs(d1, d1u, d2, d2u, sle1.p, sle2.p);
        return o_up;
    }
    // the rest of the direction from the solution (uu, vv) of that system;
    // with `bot`, DOT is first completed from the bottom system's pending
    // reductions (same flush as UP and the direction's own mix)
    int direction_finish(double eta, const Mix* bot, double& DOT, int o_up, double* hrxz, double* hrxuzu,
                         double hrtk, double* odx, double* odxu, double& odtau, double* ody, double* odyu,
                         double* odz, double* odzu, double& odkappa, double& rpn, double& rdn, double& rgn) {
        Timer tm(&t_direction);
        double *d1 = dir1.p, *d1u = dir1u.p, *d2 = dir2.p, *d2u = dir2u.p;
        const double *rdp = r_d.p, *rpp = r_p.p, *xp = x.p, *zp = z.p, *xup = xu.p, *zup = zu.p,
                     *yup = yu.p, *cp = c_.p, *bp = b_.p, *upp = u_.p, *xfp = xf.p;
        const double tau_ = tau, rho_ = rho;
        const int nlb = n_lb, nfr = n_fr;
        const int64_t* cptr = colptr; const int32_t* crow = rowind; const double* cval = cscval;
        augmented_finish(d1u, d2u, uu.p, uuu.p, vv.p, vvu.p);
        const Mix mu_v = mix_late(uu.p, vv.p, vvu.p);
        flush();
        if (bot) DOT = mix_value(*bot);
        double UP = eta * r_g + hrtk / tau;
        UP += val(o_up);
        double DOWN = D_g - DOT;
        if (DOWN <= 1e-13) odtau = 0;
        else odtau = (UP - mix_value(mu_v)) / std::max(D_g, DOWN);
        const double dt = odtau;
        const double *up_ = uu.p, *pp_ = pp.p, *uuu_ = uuu.p, *ppu_ = ppu.p, *vv_ = vv.p, *qq_ = qq.p,
                     *vvu_ = vvu.p, *qqu_ = qqu.p;
        map(n_lb, [=] __device__(int64_t i) {
            double T = up_[i] - pp_[i] * dt;
            if (xp[i] >= zp[i]) { odx[i] = T; odz[i] = hrxz[i] - zp[i] * T / xp[i]; }
            else { odx[i] = hrxz[i] + T; odz[i] = -zp[i] * T / xp[i]; }
        });
        map(n_fr, [=] __device__(int64_t i) {
            odx[i + nlb] = up_[i + nlb] - pp_[i + nlb] * dt;
            odz[i + nlb] = hrxz[i + nlb];
        });
        map(n_ub, [=] __device__(int64_t i) {
            double T = uuu_[i] - ppu_[i] * dt;
            if (xup[i] >= zup[i]) { odxu[i] = T; odzu[i] = hrxuzu[i] - zup[i] * T / xup[i]; }
            else { odxu[i] = hrxuzu[i] + T; odzu[i] = -zup[i] * T / xup[i]; }
        });
        map(m, [=] __device__(int64_t i) { ody[i] = vv_[i] - qq_[i] * dt; });
        map(n_ub, [=] __device__(int64_t i) { odyu[i] = vvu_[i] - qqu_[i] * dt; });
        odkappa = (hrtk - kappa * dt) / tau;
        // dir1 = eta r_d - c dt + dz (free: HSDSolver.cpp:435 as written), then += A^T dy
        const double rt = rho / tau;
        map(n, [=] __device__(int64_t j) {
            double acc = eta * rdp[j];
            if (j < nlb) acc += -cp[j] * dt + odz[j];
            else if (j < nlb + nfr) {
                int64_t i = j - nlb;
                acc += -(cp[i] - rt * xfp[i] * xp[j]) * dt - rho_ * xfp[i] * odx[j];
            }
            d1[j] = col_dot_acc(cptr, crow, cval, ody, j, acc);
        });
        // dir2 = eta r_p + b dt - A dx
        {
            const double* adx = tmpm.p;
            launch_csr_split(h->rows, rowptr, colind, csrval, odx, tmpm.p, false, st);
            map(m, [=] __device__(int64_t r) { d2[r] = (eta * rpp[r] + bp[r] * dt) - adx[r]; });
        }
        map(n_ub, [=] __device__(int64_t i) {
            d1[i] += odyu[i];
            d1u[i] = eta * (zup[i] + yup[i]) + odyu[i] + odzu[i];
            d2u[i] = eta * (tau_ * upp[i] - xp[i] - xup[i]) + upp[i] * dt - odxu[i] - odx[i];
        });
        const int o_d2 = dot_late(m, d2, d2), o_d2u = dot_late(n_ub, d2u, d2u);
        const int o_d1 = dot_late(n, d1, d1), o_d1u = dot_late(n_ub, d1u, d1u);
        int o_tmp = -1;
        if (n_fr > 0)
            o_tmp = reduce_late<1>(n_fr, [=] __device__(int64_t i, double* v) {
                v[0] = xfp[i] * xp[nlb + i] * xp[nlb + i];
            });
        const Mix mo = mix_late(odx, ody, odyu);
        flush();
        rpn = val(o_d2) + val(o_d2u);
        rdn = val(o_d1) + val(o_d1u);
        double tmp = 0;
        if (n_fr > 0) tmp = val(o_tmp);
        rgn = eta * r_g + odkappa - mix_value(mo) - rho * tmp * dt / (tau * tau);
        rgn = rgn * rgn;
        if (rpn + rdn <= 1e-10 || rpn + rdn + rgn <= sum_r_norm * 0.1) return 1;
        return 0;
    }

    // HSDSolver.cpp:478-488: the smallest ratio -v/dv below 1 over dv < -tol
    // (stepsize_late launches the reductions, stepsize_value reads them)
    std::pair<int, int> stepsize_late(const double* v, const double* vu, const double* dv, const double* dvu) {
        const int a = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* o) {
            o[0] = (dv[i] < -kVarTol && v[i] + dv[i] < 0.0) ? -v[i] / dv[i] : HUGE_VAL;
        });
        int b = -1;
        if (n_ub > 0)
            b = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* o) {
                o[0] = (dvu[i] < -kVarTol && vu[i] + dvu[i] < 0.0) ? -vu[i] / dvu[i] : HUGE_VAL;
            });
        return {a, b};
    }
    double stepsize_value(std::pair<int, int> o) const {
        double r = 1.0;
        r = std::min(r, val(o.first, 1));
        if (o.second >= 0) r = std::min(r, val(o.second, 1));
        return r;
    }

    // HSDSolver.cpp:491-514
    int check_step(double ap, double at, double ad, double ak, double beta) {
        double xz_min = (tau + at * dtau) * (kappa + ak * dkappa);
        double xz_sum = xz_min;
        const double *xp = x.p, *dxp = dx.p, *zp = z.p, *dzp = dz.p;
        const int a = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* o) {
            o[0] = (xp[i] + ap * dxp[i]) * (zp[i] + ad * dzp[i]);
        });
        int b = -1;
        if (n_ub > 0) {
            const double *xup = xu.p, *dxup = dxu.p, *zup = zu.p, *dzup = dzu.p;
            b = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* o) {
                o[0] = (xup[i] + ap * dxup[i]) * (zup[i] + ad * dzup[i]);
            });
        }
        flush();
        xz_min = std::min(xz_min, val(a, 1));
        xz_sum += val(a);
        if (b >= 0) {
            xz_min = std::min(xz_min, val(b, 1));
            xz_sum += val(b);
        }
        double mu_ = xz_sum / (1.0 + n_lb + n_ub);
        return xz_min / mu_ >= beta;
    }

    void get_stepsize(double& ps, double& ds) {
        Timer tm(&t_step);
        const auto op = stepsize_late(x.p, xu.p, dx.p, dxu.p);
        const auto od = stepsize_late(z.p, zu.p, dz.p, dzu.p);
        flush();
        ps = stepsize_value(op);
        ds = stepsize_value(od);
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
        double *xp = x.p, *zp = z.p, *yp = y.p, *xup = xu.p, *yup = yu.p, *zup = zu.p;
        const double *dxp = dx.p, *dzp = dz.p, *dyp = dy.p, *dxup = dxu.p, *dyup = dyu.p, *dzup = dzu.p;
        const int nlb = n_lb;
        map(n_lb, [=] __device__(int64_t i) {
            xp[i] = (xp[i] + p_step * dxp[i]) * pf;
            zp[i] = (zp[i] + d_step * dzp[i]) * df;
        });
        map(m, [=] __device__(int64_t i) { yp[i] = (yp[i] + d_step * dyp[i]) * df; });
        map(n_ub, [=] __device__(int64_t i) {
            xup[i] = (xup[i] + p_step * dxup[i]) * pf;
            yup[i] = (yup[i] + d_step * dyup[i]) * df;
            zup[i] = (zup[i] + d_step * dzup[i]) * df;
        });
        map(n_fr, [=] __device__(int64_t i) { xp[i + nlb] = (xp[i + nlb] + p_step * dxp[i + nlb]) * pf; });
        tau += tau_step * dtau;
        kappa = (kappa + kappa_step * dkappa) * pf * df;
    }

    void take_improved(double dtau_b, double dkappa_b) {
        auto cp = [&](DBuf<double>& d, DBuf<double>& s, int k) {
            if (k > 0) HIPCHK(hipMemcpyAsync(d.p, s.p, sizeof(double) * k, hipMemcpyDeviceToDevice, st));
        };
        cp(dx, dx_b, n); cp(dxu, dxu_b, n_ub); cp(dy, dy_b, m); cp(dyu, dyu_b, n_ub); cp(dz, dz_b, n);
        cp(dzu, dzu_b, n_ub);
        dtau = dtau_b;
        dkappa = dkappa_b;
    }

    // hat r_xz = -x z + alpha - phi (dx dz - dmu) (and the upper-bound / tau-kappa parts)
    void set_hat(double alpha, double phi, double dmu, bool with_dir) {
        double *hx = hxz.p, *hxu = hxuzu.p;
        const double *xp = x.p, *zp = z.p, *xup = xu.p, *zup = zu.p, *dxp = dx.p, *dzp = dz.p, *dxup = dxu.p,
                     *dzup = dzu.p;
        const int nlb = n_lb;
        if (with_dir) {
            map(n_lb, [=] __device__(int64_t i) { hx[i] = -xp[i] * zp[i] + alpha - phi * (dxp[i] * dzp[i] - dmu); });
            map(n_ub, [=] __device__(int64_t i) {
                hxu[i] = -xup[i] * zup[i] + alpha - phi * (dxup[i] * dzup[i] - dmu);
            });
            htk = -tau * kappa + alpha - phi * (dtau * dkappa - dmu);
        } else {
            map(n_lb, [=] __device__(int64_t i) { hx[i] = -xp[i] * zp[i] + alpha; });
            map(n_ub, [=] __device__(int64_t i) { hxu[i] = -xup[i] * zup[i] + alpha; });
            htk = -tau * kappa + alpha;
        }
        map(n_fr, [=] __device__(int64_t i) { hx[i + nlb] = 0; });
    }

    double dmu_of(double nall) {
        const int a = dot_late(n_lb, dx.p, dz.p), b = dot_late(n_ub, dxu.p, dzu.p);
        flush();
        return std::max((val(a) + val(b) + dtau * dkappa) / nall, 0.0);
    }

    // HSDSolver.cpp:596-717
    void init_point() {
        status = HELP_HSD_OK;
        double *xp = x.p, *zp = z.p, *xup = xu.p, *zup = zu.p, *yup = yu.p, *yp = y.p, *xfp = xf.p;
        const int nlb = n_lb;
        map(n_lb, [=] __device__(int64_t i) { xp[i] = 1.0; zp[i] = 1.0; });
        map(n_ub, [=] __device__(int64_t i) { xup[i] = 1.0; zup[i] = 1.0; yup[i] = 0.0; });
        map(m, [=] __device__(int64_t i) { yp[i] = 0.0; });
        tau = kappa = 1.0;
        map(n_fr, [=] __device__(int64_t i) {
            xp[nlb + i] = 0.0; zp[nlb + i] = 0.0;
            xfp[i] = 1.0 / (1.0 + 10.0 * fabs(xp[nlb + i]));
        });
        double nall = 1.0 + n_lb + n_ub;
        double mu0 = (dot(n_lb, xp, zp) + dot(n_ub, xup, zup) + tau * kappa) / nall;
        inf0 = tau / kappa;
        rho = 1.0;
        newton_parameters();
        if (status != HELP_HSD_OK) return;
        renew();
        bottom_rhs();
        double DOT, rpn, rdn, rgn, phi = 1.0, eta = 1.0;
        set_hat(mu0, 0.0, 0.0, false);
        int newton_ok = bottom_and_direction(eta, DOT, hxz.p, hxuzu.p, htk, dx.p, dxu.p, dtau, dy.p, dyu.p, dz.p,
                                             dzu.p, dkappa, rpn, rdn, rgn);
        double p_step = 0, d_step = 0;
        if (newton_ok) get_stepsize(p_step, d_step);
        else phi = 0.0;
        double gamma = 10.0;
        double alpha_p = std::max(0.0, 1.0 - std::min(p_step, d_step)) * gamma * mu0;
        double dmu = dmu_of(nall);
        set_hat(alpha_p, phi, dmu, true);
        double dtau_b, dkappa_b, rpb, rdb, rgb;
        int imp_ok = direction(eta, DOT, hxz.p, hxuzu.p, htk, dx_b.p, dxu_b.p, dtau_b, dy_b.p, dyu_b.p, dz_b.p,
                               dzu_b.p, dkappa_b, rpb, rdb, rgb);
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
        double nall = 1.0 + n_lb + n_ub;
        double mu0 = (dot(n_lb, x.p, z.p) + dot(n_ub, xu.p, zu.p) + tau * kappa) / nall;
        mu = mu0;
        int it;
        const double *xp = x.p;
        double* xfp = xf.p;
        const int nlb = n_lb;
        for (it = 0; it < max_iter; ++it) {
            // a failed backend call (e.g. PCG stopped at its cap short of
            // its tolerance) ends the run: the reference would exit in
            // CheckError (Helper.cpp:24-35)
            if (backend_error) { status = HELP_HSD_STALLED; stall_reason = HELP_STALL_BACKEND; break; }
            rho = std::max(mu / mu0, 1e-12);
            map(n_fr, [=] __device__(int64_t i) { xfp[i] = 1.0 / (1.0 + 10.0 * fabs(xp[nlb + i])); });
            newton_parameters();
            if (logf) logf(user, it, pobj, dobj, rel_gap, pinf, dinf);
            if (status != HELP_HSD_OK || (mu <= kMuTol * mu0 && infe < kInfeasTol) ||
                (rel_gap <= kGapTol && pinf <= kPinfTol && dinf <= kDinfTol))
                break;
            renew();
            bottom_rhs();
            double DOT, rpn, rdn, rgn, phi = 1.0, eta = 1.0;
            set_hat(0.0, 0.0, 0.0, false);
            int newton_ok = bottom_and_direction(eta, DOT, hxz.p, hxuzu.p, htk, dx.p, dxu.p, dtau, dy.p, dyu.p,
                                                 dz.p, dzu.p, dkappa, rpn, rdn, rgn);
            double p_step = 0, d_step = 0, gamma;
            int o_dmu[2] = {-1, -1};
            if (newton_ok) {
                get_stepsize(p_step, d_step);
                const double *dxp = dx.p, *zp = z.p, *dzp = dz.p;
                const int a = reduce_late<1>(n_lb, [=] __device__(int64_t i, double* o) {
                    o[0] = (xp[i] + p_step * dxp[i]) * (zp[i] + d_step * dzp[i]);
                });
                int b = -1;
                if (n_ub > 0) {
                    const double *xup = xu.p, *dxup = dxu.p, *zup = zu.p, *dzup = dzu.p;
                    b = reduce_late<1>(n_ub, [=] __device__(int64_t i, double* o) {
                        o[0] = (xup[i] + p_step * dxup[i]) * (zup[i] + d_step * dzup[i]);
                    });
                }
                // the dmu dot products (read below) share this flush
                o_dmu[0] = dot_late(n_lb, dx.p, dz.p);
                o_dmu[1] = dot_late(n_ub, dxu.p, dzu.p);
                flush();
                double new_mu = (tau + p_step * dtau) * (kappa + d_step * dkappa);
                new_mu += val(a);
                if (b >= 0) new_mu += val(b);
                new_mu /= nall;
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
            double dmu = std::max((val(o_dmu[0]) + val(o_dmu[1]) + dtau * dkappa) / nall, 0.0);   // dmu_of
            set_hat(alpha_p, phi, dmu, true);
            double dtau_b, dkappa_b, rpb, rdb, rgb;
            int imp_ok = direction(eta, DOT, hxz.p, hxuzu.p, htk, dx_b.p, dxu_b.p, dtau_b, dy_b.p, dyu_b.p,
                                   dz_b.p, dzu_b.p, dkappa_b, rpb, rdb, rgb);
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
}  // namespace help

extern "C" int help_hsd_solve_gpu(const help_lp* lp, const help_backend* be, int32_t max_iter,
                                  help_hsd_log_fn logf, void* user, help_hsd_result* res) {
    using namespace help;
    if (!lp || !be || !res || !be->construct || !be->destruct) return HELP_ERR_ARG;
    if (lp->n_lb + lp->n_fr != lp->n || lp->n_ub > lp->n_lb || (lp->n_ub > 0 && !lp->ub)) return HELP_ERR_ARG;
    std::memset(res, 0, sizeof(*res));
    if (max_iter <= 0) max_iter = 10000;
    double t0 = now_s();
    int st = be->construct(be->ctx, lp);
    res->seconds_construct = now_s() - t0;
    help_gpu* h = st ? nullptr : help_gpu_backend_handle(be);
    if (st || !h || !h->device) {
        res->backend_error = st ? st : HELP_ERR_STATE;
        be->destruct(be->ctx);
        res->seconds = now_s() - t0;
        return res->backend_error;
    }
    int it = 0;
    try {
        HsdGpu g(lp, be, h);
        it = g.solve_all(max_iter, logf, user);
        res->status = g.status;
        res->stall_reason = g.stall_reason;
        res->newton_ratio = g.newton_ratio;
        res->primal_obj = g.pobj;
        res->dual_obj = g.dobj;
        res->rel_gap = g.rel_gap;
        res->primal_infeas = g.pinf;
        res->dual_infeas = g.dinf;
        res->mu = g.mu;
        res->seconds_renew = g.t_renew;
        res->seconds_solve = g.t_solve;
        res->seconds_newton = g.t_newton;
        res->seconds_direction = g.t_direction;
        res->seconds_step = g.t_step;
        res->n_renew = g.n_renew;
        res->n_solve = g.n_solve;
        res->backend_error = g.backend_error;
    } catch (const std::exception& e) {
        h->err = e.what();
        res->backend_error = HELP_ERR_HIP;
    }
    be->destruct(be->ctx);
    res->iterations = it;
    res->seconds = now_s() - t0;
    return res->backend_error == HELP_ERR_HIP ? HELP_ERR_HIP : HELP_OK;
}
