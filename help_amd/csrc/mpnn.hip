// Host side of the warm-start network: parameters (FP32 master, the copy
// the kernels use, Adam moments), the per-instance workspace, the forward
// and backward passes over the handle's matrix, and the C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "handle.h"
#include "mpnn.h"

using help::Bf16;
using help::NnLayout;
using help::kNnH;

struct help_mpnn {
    int bf16 = 1;
    int step = 0;
    DBuf<float> p, pw, grad, mo, ve;
    // workspace for the largest instance seen
    int64_t cap_m = 0, cap_n = 0;
    DBuf<double> dd, dr, ddy, dinv, g, x0, stat;
    DBuf<float> y, dyv, xr, xc, R1, R2, C1, C2, P, Q, partial;
    DBuf<uint8_t> act;
    double rs = 1, mu = 0, nu = 0, loss = 0;
    bool have_grad = false;
    int32_t last_m = 0, last_n = 0;
};

namespace {

constexpr int kStatBlocks = 256;

float host_round_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

void upload_params(help_mpnn* net, const std::vector<float>& p) {
    std::vector<float> pw(p.size());
    for (size_t i = 0; i < p.size(); ++i) pw[i] = net->bf16 ? host_round_bf16(p[i]) : p[i];
    HIPCHK(hipMemcpy(net->p.p, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(net->pw.p, pw.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
}

void ensure_workspace(help_mpnn* net, int32_t m, int32_t n) {
    if (m <= net->cap_m && n <= net->cap_n) return;
    const int64_t M = std::max<int64_t>(m, net->cap_m), N = std::max<int64_t>(n, net->cap_n);
    net->cap_m = M;
    net->cap_n = N;
    const int64_t MN = std::max(M, N);
    net->dd.alloc(N); net->dinv.alloc(N); net->dr.alloc(M); net->ddy.alloc(M); net->g.alloc(M); net->x0.alloc(M);
    net->stat.alloc(2 * kStatBlocks);
    net->y.alloc(M); net->dyv.alloc(M);
    net->xr.alloc(M * help::kNnFR); net->xc.alloc(N * help::kNnFC);
    net->R1.alloc(M * kNnH); net->R2.alloc(M * kNnH);
    net->C1.alloc(N * kNnH); net->C2.alloc(N * kNnH);
    net->P.alloc(MN * kNnH); net->Q.alloc(MN * kNnH);
    const int64_t nblk = (MN + help::kNnChunk - 1) / help::kNnChunk;
    net->partial.alloc(nblk * (2 * kNnH + 1) * kNnH);
    const size_t esz = net->bf16 ? sizeof(Bf16) : sizeof(float);
    net->act.alloc((size_t)kNnH * esz * (6 * M + 5 * N));
}

// activation buffers carved out of net->act
template <typename AT>
struct Acts {
    AT *hr0, *ar2, *hr1, *ar4, *hr2, *z, *hc0, *ac1, *hc1, *ac3, *hc2;
    Acts(help_mpnn* net) {
        AT* base = (AT*)net->act.p;
        const int64_t M = net->cap_m * kNnH, N = net->cap_n * kNnH;
        hr0 = base; ar2 = hr0 + M; hr1 = ar2 + M; ar4 = hr1 + M; hr2 = ar4 + M; z = hr2 + M;
        hc0 = z + M; ac1 = hc0 + N; hc1 = ac1 + N; ac3 = hc1 + N; hc2 = ac3 + N;
    }
};

double host_sum(help_mpnn* net, hipStream_t st, int64_t n, const double* x, int mode, double* count) {
    help::nn_logsum_partials(n, x, mode, net->stat.p, kStatBlocks, st);
    std::vector<double> h(2 * kStatBlocks);
    HIPCHK(hipMemcpyAsync(h.data(), net->stat.p, sizeof(double) * (mode == 0 ? 2 : 1) * kStatBlocks,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double s = 0, c = 0;
    for (int b = 0; b < kStatBlocks; ++b) s += h[b];
    if (mode == 0)
        for (int b = 0; b < kStatBlocks; ++b) c += h[kStatBlocks + b];
    if (count) *count = c;
    return s;
}

// forward pass; d, r device vectors of the handle's size; x0 out (device)
template <typename AT>
void forward(help_mpnn* net, help_gpu* h, const double* d, const double* r, double* x0) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    ensure_workspace(net, m, n);
    net->last_m = m;
    net->last_n = n;
    help::launch_dinv(d, net->dinv.p, n, st);
    help::launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, net->dinv.p, net->g.p, true, st);
    double cnt = 0;
    const double lg = host_sum(net, st, m, net->g.p, 0, &cnt);
    net->mu = cnt > 0 ? lg / cnt : 0.0;
    const double ld = host_sum(net, st, n, net->dinv.p, 0, &cnt);
    net->nu = cnt > 0 ? ld / cnt : 0.0;
    const double r2 = host_sum(net, st, m, r, 1, nullptr);
    net->rs = r2 > 0 ? std::sqrt(r2 / m) : 1.0;
    help::nn_features(m, n, net->g.p, r, net->dinv.p, net->rs, net->mu, net->nu, net->xr.p, net->xc.p, st);
    Acts<AT> a(net);
    const float* W = net->pw.p;
    const int64_t *rp = h->d_rowptr.p, *cp = h->d_colptr.p;
    const int32_t *ci = h->d_colind.p, *ri = h->d_rowind.p;
    const double *rv = h->d_csrval.p, *cv = h->d_val.p;
    help::nn_dense<AT, float>(m, net->xr.p, help::kNnFR, nullptr, 0, W + NnLayout::Wr0, W + NnLayout::br0, a.hr0, st);
    help::nn_dense<AT, float>(n, net->xc.p, help::kNnFC, nullptr, 0, W + NnLayout::Wc0, W + NnLayout::bc0, a.hc0, st);
    help::nn_mean<AT>(n, cp, ri, cv, a.hr0, a.ac1, st);                                     // L1
    help::nn_dense<AT, AT>(n, a.hc0, kNnH, a.ac1, kNnH, W + NnLayout::W(1), W + NnLayout::b(1), a.hc1, st);
    help::nn_mean<AT>(m, rp, ci, rv, a.hc1, a.ar2, st);                                     // L2
    help::nn_dense<AT, AT>(m, a.hr0, kNnH, a.ar2, kNnH, W + NnLayout::W(2), W + NnLayout::b(2), a.hr1, st);
    help::nn_mean<AT>(n, cp, ri, cv, a.hr1, a.ac3, st);                                     // L3
    help::nn_dense<AT, AT>(n, a.hc1, kNnH, a.ac3, kNnH, W + NnLayout::W(3), W + NnLayout::b(3), a.hc2, st);
    help::nn_mean<AT>(m, rp, ci, rv, a.hc2, a.ar4, st);                                     // L4
    help::nn_dense<AT, AT>(m, a.hr1, kNnH, a.ar4, kNnH, W + NnLayout::W(4), W + NnLayout::b(4), a.hr2, st);
    help::nn_dense<AT, AT>(m, a.hr2, kNnH, nullptr, 0, W + NnLayout::Wh, W + NnLayout::bh, a.z, st);
    help::nn_output<AT>(m, a.z, W + NnLayout::wo, W + NnLayout::bo, net->g.p, net->rs, net->y.p, x0, st);
    HIPCHK(hipGetLastError());
}

// weight gradient of one dense layer into grad[off .. off + (k1+k2+1)*64)
template <typename AT, typename IT>
void dense_wgrad(help_mpnn* net, hipStream_t st, int64_t N, const IT* in1, int k1, const AT* in2, int k2,
                 const float* dpre, int off) {
    const int nblk = (int)((N + help::kNnChunk - 1) / help::kNnChunk);
    help::nn_dense_bwd_w<AT, IT>(N, in1, k1, in2, k2, dpre, net->partial.p, st);
    help::nn_reduce_blocks((k1 + k2 + 1) * kNnH, nblk, net->partial.p, net->grad.p + off, st);
}

template <typename AT>
void backward(help_mpnn* net, help_gpu* h, const double* dy) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    Acts<AT> a(net);
    const float* W = net->pw.p;
    const int64_t *rp = h->d_rowptr.p, *cp = h->d_colptr.p;
    const int32_t *ci = h->d_colind.p, *ri = h->d_rowind.p;
    const double *rv = h->d_csrval.p, *cv = h->d_val.p;
    // loss = sum (y - t)^2 / sum t^2
    help::nn_loss_partials(m, net->y.p, dy, net->g.p, net->rs, net->stat.p, kStatBlocks, st);
    std::vector<double> hs(2 * kStatBlocks);
    HIPCHK(hipMemcpyAsync(hs.data(), net->stat.p, sizeof(double) * 2 * kStatBlocks, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double E = 0, T = 0;
    for (int b = 0; b < kStatBlocks; ++b) { E += hs[b]; T += hs[kStatBlocks + b]; }
    if (!(T > 0)) T = 1.0;
    net->loss = E / T;
    help::nn_output_grad(m, net->y.p, dy, net->g.p, net->rs, T, net->dyv.p, st);
    float *R1 = net->R1.p, *R2 = net->R2.p, *C1 = net->C1.p, *C2 = net->C2.p, *P = net->P.p, *Q = net->Q.p;
    // head
    help::nn_out_w<AT>(m, a.z, net->dyv.p, net->partial.p, st);
    help::nn_reduce_blocks(kNnH + 1, (m + help::kNnChunk - 1) / help::kNnChunk, net->partial.p,
                           net->grad.p + NnLayout::wo, st);
    help::nn_out_bwd<AT>(m, a.z, net->dyv.p, W + NnLayout::wo, P, st);
    dense_wgrad<AT, AT>(net, st, m, a.hr2, kNnH, (const AT*)nullptr, 0, P, NnLayout::Wh);
    help::nn_dense_bwd_in(m, P, W + NnLayout::Wh, kNnH, 0, R1, nullptr, false, st);            // d hr2
    // L4: hr2 = relu([hr1 ar4] W4 + b4), ar4 = mean(hc2)
    help::nn_relu_bwd<AT>(m, a.hr2, R1, P, st);
    dense_wgrad<AT, AT>(net, st, m, a.hr1, kNnH, a.ar4, kNnH, P, NnLayout::W(4));
    help::nn_dense_bwd_in(m, P, W + NnLayout::W(4), kNnH, kNnH, R2, Q, false, st);               // d hr1, d ar4
    HIPCHK(hipMemsetAsync(C1, 0, sizeof(float) * (size_t)n * kNnH, st));
    help::nn_mean_bwd(n, cp, ri, cv, rp, Q, C1, st);                                            // d hc2
    // L3: hc2 = relu([hc1 ac3] W3 + b3), ac3 = meanT(hr1)
    help::nn_relu_bwd<AT>(n, a.hc2, C1, P, st);
    dense_wgrad<AT, AT>(net, st, n, a.hc1, kNnH, a.ac3, kNnH, P, NnLayout::W(3));
    help::nn_dense_bwd_in(n, P, W + NnLayout::W(3), kNnH, kNnH, C2, Q, false, st);               // d hc1, d ac3
    help::nn_mean_bwd(m, rp, ci, rv, cp, Q, R2, st);                                            // d hr1 +=
    // L2: hr1 = relu([hr0 ar2] W2 + b2), ar2 = mean(hc1)
    help::nn_relu_bwd<AT>(m, a.hr1, R2, P, st);
    dense_wgrad<AT, AT>(net, st, m, a.hr0, kNnH, a.ar2, kNnH, P, NnLayout::W(2));
    help::nn_dense_bwd_in(m, P, W + NnLayout::W(2), kNnH, kNnH, R1, Q, false, st);               // d hr0, d ar2
    help::nn_mean_bwd(n, cp, ri, cv, rp, Q, C2, st);                                            // d hc1 +=
    // L1: hc1 = relu([hc0 ac1] W1 + b1), ac1 = meanT(hr0)
    help::nn_relu_bwd<AT>(n, a.hc1, C2, P, st);
    dense_wgrad<AT, AT>(net, st, n, a.hc0, kNnH, a.ac1, kNnH, P, NnLayout::W(1));
    help::nn_dense_bwd_in(n, P, W + NnLayout::W(1), kNnH, kNnH, C1, Q, false, st);               // d hc0, d ac1
    help::nn_mean_bwd(m, rp, ci, rv, cp, Q, R1, st);                                            // d hr0 +=
    // input layers
    help::nn_relu_bwd<AT>(m, a.hr0, R1, P, st);
    dense_wgrad<AT, float>(net, st, m, net->xr.p, help::kNnFR, (const AT*)nullptr, 0, P, NnLayout::Wr0);
    help::nn_relu_bwd<AT>(n, a.hc0, C1, P, st);
    dense_wgrad<AT, float>(net, st, n, net->xc.p, help::kNnFC, (const AT*)nullptr, 0, P, NnLayout::Wc0);
    HIPCHK(hipGetLastError());
    net->have_grad = true;
}

int fail_msg(help_gpu* h, const std::exception& e) {
    if (h) h->err = e.what();
    if (auto* he = dynamic_cast<const HipError*>(&e)) return he->code;
    return HELP_ERR_ARG;
}

}  // namespace

namespace help {
// x0 (device) = net(A, d, r) with device d, r; used by the learned PCG mode
void mpnn_predict_device(help_mpnn* net, help_gpu* h, const double* d, const double* r, double* x0) {
    if (net->bf16) forward<Bf16>(net, h, d, r, x0);
    else forward<float>(net, h, d, r, x0);
}
}  // namespace help

extern "C" {

int help_mpnn_num_params(void) { return NnLayout::total; }

int help_selftest_mfma_bf16(const double* a, const double* b, double* out) {
    if (!a || !b || !out) return HELP_ERR_ARG;
    try {
        if (help_device_count() == 0) return HELP_ERR_NOGPU;
        std::vector<float> fa(a, a + 512), fb(b, b + 512), fo(256);
        DBuf<float> da, db, dout;
        da.alloc(512); db.alloc(512); dout.alloc(256);
        HIPCHK(hipMemcpy(da.p, fa.data(), 512 * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(db.p, fb.data(), 512 * sizeof(float), hipMemcpyHostToDevice));
        help::nn_mfma_bf16_probe(da.p, db.p, dout.p, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(fo.data(), dout.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
        for (int i = 0; i < 256; ++i) out[i] = fo[i];
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(nullptr, e);
    }
}

int help_mpnn_new(int32_t bf16, uint64_t seed, help_mpnn** out) {
    if (!out) return HELP_ERR_ARG;
    *out = nullptr;
    try {
        if (help_device_count() == 0) return HELP_ERR_NOGPU;
        auto* net = new help_mpnn();
        net->bf16 = bf16 ? 1 : 0;
        const int np = NnLayout::total;
        net->p.alloc(np); net->pw.alloc(np); net->grad.alloc(np); net->mo.alloc(np); net->ve.alloc(np);
        HIPCHK(hipMemset(net->mo.p, 0, sizeof(float) * np));
        HIPCHK(hipMemset(net->ve.p, 0, sizeof(float) * np));
        HIPCHK(hipMemset(net->grad.p, 0, sizeof(float) * np));
        // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, a
        // small output layer (the first guesses are close to zero)
        std::vector<float> p(np, 0.0f);
        uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        auto unif = [&]() {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        auto fill = [&](int off, int fan_in, int count, double gain) {
            const double a = gain / std::sqrt((double)fan_in);
            for (int i = 0; i < count; ++i) p[off + i] = (float)(a * unif());
        };
        fill(NnLayout::Wr0, help::kNnFR, help::kNnFR * kNnH, 1.0);
        fill(NnLayout::Wc0, help::kNnFC, help::kNnFC * kNnH, 1.0);
        for (int l = 1; l <= 4; ++l) fill(NnLayout::W(l), 2 * kNnH, 2 * kNnH * kNnH, 1.0);
        fill(NnLayout::Wh, kNnH, kNnH * kNnH, 1.0);
        fill(NnLayout::wo, kNnH, kNnH, 0.1);
        upload_params(net, p);
        *out = net;
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(nullptr, e);
    }
}

void help_mpnn_free(help_mpnn* net) { delete net; }

int help_mpnn_get_params(const help_mpnn* net, float* out) {
    if (!net || !out) return HELP_ERR_ARG;
    try {
        HIPCHK(hipMemcpy(out, net->p.p, sizeof(float) * NnLayout::total, hipMemcpyDeviceToHost));
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(nullptr, e);
    }
}

int help_mpnn_set_params(help_mpnn* net, const float* in) {
    if (!net || !in) return HELP_ERR_ARG;
    try {
        upload_params(net, std::vector<float>(in, in + NnLayout::total));
        HIPCHK(hipMemset(net->mo.p, 0, sizeof(float) * NnLayout::total));
        HIPCHK(hipMemset(net->ve.p, 0, sizeof(float) * NnLayout::total));
        net->step = 0;
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(nullptr, e);
    }
}

int help_mpnn_predict(help_mpnn* net, help_gpu* h, const double* d, const double* r, double* x0) {
    if (!net || !h || !d || !r || !x0) return HELP_ERR_ARG;
    if (!h->device) { h->err = "mpnn_predict: handle has no device state"; return HELP_ERR_STATE; }
    try {
        ensure_workspace(net, h->m, h->n);
        hipStream_t st = h->st;
        HIPCHK(hipMemcpyAsync(net->dd.p, d, sizeof(double) * h->n, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(net->dr.p, r, sizeof(double) * h->m, hipMemcpyDefault, st));
        help::mpnn_predict_device(net, h, net->dd.p, net->dr.p, net->x0.p);
        HIPCHK(hipMemcpyAsync(x0, net->x0.p, sizeof(double) * h->m, hipMemcpyDefault, st));
        HIPCHK(hipStreamSynchronize(st));
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(h, e);
    }
}

int help_mpnn_loss_grad(help_mpnn* net, help_gpu* h, const double* d, const double* r, const double* dy,
                        double* loss, float* grad) {
    if (!net || !h || !d || !r || !dy) return HELP_ERR_ARG;
    if (!h->device) { h->err = "mpnn_loss_grad: handle has no device state"; return HELP_ERR_STATE; }
    try {
        ensure_workspace(net, h->m, h->n);
        hipStream_t st = h->st;
        HIPCHK(hipMemcpyAsync(net->dd.p, d, sizeof(double) * h->n, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(net->dr.p, r, sizeof(double) * h->m, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(net->ddy.p, dy, sizeof(double) * h->m, hipMemcpyDefault, st));
        if (net->bf16) {
            forward<Bf16>(net, h, net->dd.p, net->dr.p, nullptr);
            backward<Bf16>(net, h, net->ddy.p);
        } else {
            forward<float>(net, h, net->dd.p, net->dr.p, nullptr);
            backward<float>(net, h, net->ddy.p);
        }
        if (grad) HIPCHK(hipMemcpyAsync(grad, net->grad.p, sizeof(float) * NnLayout::total, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (loss) *loss = net->loss;
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(h, e);
    }
}

int help_mpnn_adam_step(help_mpnn* net, double lr) {
    if (!net || !net->have_grad) return HELP_ERR_STATE;
    try {
        net->step += 1;
        help::nn_adam(NnLayout::total, net->p.p, net->grad.p, net->mo.p, net->ve.p, (float)lr, 0.9f, 0.999f, 1e-8f,
                      net->step, net->bf16, net->pw.p, nullptr);
        HIPCHK(hipDeviceSynchronize());
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail_msg(nullptr, e);
    }
}

}  // extern "C"
