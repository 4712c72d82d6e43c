// Kernels of the warm-start network (the math is stated in mpnn.h): input
// features, dense layers (LDS-staged weights, FP32 accumulation), mean
// aggregation over A's rows or columns (a wave per node, features on the
// lanes), the output head, the loss, and the backward pass: ReLU masks,
// input gradients, weight gradients as fixed-order block partials, the
// transposed aggregation, and Adam.  Nothing uses atomics: every sum has a
// fixed order, so training is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "launch.h"
#include "mpnn.h"

namespace help {

namespace {

constexpr int kB = 256;
inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kB - 1) / kB); }
inline int grid_cap(int64_t n, int cap) { return (int)std::min<int64_t>(cap, std::max<int64_t>(1, n)); }

__device__ inline float ldv(float x) { return x; }
__device__ inline float ldv(Bf16 x) { return __uint_as_float((uint32_t)x.v << 16); }
__device__ inline void stv(float& d, float x) { d = x; }
__device__ inline void stv(Bf16& d, float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) { d.v = 0x7fc0; return; }
    u += 0x7fffu + ((u >> 16) & 1u);
    d.v = (uint16_t)(u >> 16);
}
__device__ inline float round_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ inline float wave_sum_f(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ---------------------------------------------------------------- features
__global__ void k_feat_row(int32_t m, const double* __restrict__ g, const double* __restrict__ r, double rs,
                           double mu, float* __restrict__ xr) {
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const double gi = g[i] > 0 ? g[i] : 1.0;
    const double jac = fmin(10.0, fmax(-10.0, exp(mu) * r[i] / (rs * gi)));
    xr[(int64_t)i * kNnFR + 0] = (float)(r[i] / rs);
    xr[(int64_t)i * kNnFR + 1] = (float)((log(gi) - mu) / 10.0);
    xr[(int64_t)i * kNnFR + 2] = (float)(jac / 10.0);
    xr[(int64_t)i * kNnFR + 3] = 1.0f;
}

__global__ void k_feat_col(int32_t n, const double* __restrict__ dinv, double nu, float* __restrict__ xc) {
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    xc[(int64_t)j * kNnFC + 0] = (float)((log(dinv[j]) - nu) / 10.0);
    xc[(int64_t)j * kNnFC + 1] = 1.0f;
}

// mode 0: part[b] = sum of log(x) over x > 0, part[nblk + b] = their count;
// mode 1: part[b] = sum of x^2
__global__ __launch_bounds__(256) void k_logsum(int64_t n, const double* __restrict__ x, int mode,
                                                double* __restrict__ part) {
    __shared__ double s0[kB], s1[kB];
    double a = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const double v = x[i];
        if (mode == 0) {
            if (v > 0) { a += log(v); c += 1.0; }
        } else {
            a += v * v;
        }
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s0[threadIdx.x] += s0[threadIdx.x + w]; s1[threadIdx.x] += s1[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0[0];
        if (mode == 0) part[gridDim.x + blockIdx.x] = s1[0];
    }
}

// ------------------------------------------------------------ dense layers
template <typename AT, typename IT>
__global__ __launch_bounds__(256) void k_dense(int64_t N, const IT* __restrict__ in1, int k1,
                                               const AT* __restrict__ in2, int k2, const float* __restrict__ W,
                                               const float* __restrict__ b, AT* __restrict__ out) {
    __shared__ float Ws[2 * kNnH * kNnH];
    __shared__ float xs[4][2 * kNnH];
    const int K = k1 + k2;
    for (int t = threadIdx.x; t < K * kNnH; t += kB) Ws[t] = W[t];
    const int q = threadIdx.x >> 6, o = threadIdx.x & 63;
    const float bo = b[o];
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * 4; base < N; base += (int64_t)gridDim.x * 4) {
        for (int t = threadIdx.x; t < 4 * K; t += kB) {
            const int qq = t / K, k = t - qq * K;
            const int64_t node = base + qq;
            float v = 0.0f;
            if (node < N) v = (k < k1) ? ldv(in1[node * k1 + k]) : ldv(in2[node * k2 + (k - k1)]);
            xs[qq][k] = v;
        }
        __syncthreads();
        const int64_t node = base + q;
        if (node < N) {
            float acc = bo;
            for (int k = 0; k < K; ++k) acc += xs[q][k] * Ws[k * kNnH + o];
            stv(out[node * kNnH + o], fmaxf(acc, 0.0f));
        }
        __syncthreads();
    }
}

// The BF16 layers with K = k1 + k2 a multiple of 32 (the message layers and
// the head) run on V_MFMA_F32_16X16X32_BF16 with FP32 accumulation: a
// workgroup takes 64 nodes, each wave 16 of them across the 64 outputs (four
// 16x16 tiles, the bias as the accumulators' start).  Lane layout (checked by
// help_selftest_mfma_bf16): A[i][k] in lane i + 16 (k/8), element k%8;
// B[k][j] in lane j + 16 (k/8), element k%8; D[i][j] in lane j + 16 (i/4),
// element i%4.  A lane's 8 inputs are 16 contiguous bytes of one node's row;
// the weights are staged in LDS transposed (row o holds W[.][o], padded by 8
// elements against bank conflicts) as BF16, which is exact: the BF16 mode's
// weights are BF16-rounded.
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4_t;

__device__ inline uint16_t bf16_bits(float x) {
    Bf16 h;
    stv(h, x);
    return h.v;
}

__global__ __launch_bounds__(256) void k_dense_mfma(int64_t N, const Bf16* __restrict__ in1, int k1,
                                                    const Bf16* __restrict__ in2, int k2,
                                                    const float* __restrict__ W, const float* __restrict__ b,
                                                    Bf16* __restrict__ out) {
    constexpr int kLd = 2 * kNnH + 8;
    __shared__ __attribute__((aligned(16))) uint16_t Wt[kNnH * kLd];
    const int K = k1 + k2;
    for (int t = threadIdx.x; t < K * kNnH; t += kB) {
        const int k = t / kNnH, o = t - k * kNnH;
        Wt[o * kLd + k] = bf16_bits(W[t]);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    float bias[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) bias[t] = b[16 * t + li];
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * 64; base < N; base += (int64_t)gridDim.x * 64) {
        const int64_t row = base + wave * 16 + li;
        const int64_t rc = row < N ? row : N - 1;
        f32x4_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4_t{bias[t], bias[t], bias[t], bias[t]};
        for (int kk = 0; kk < K; kk += 32) {
            const int k0 = kk + 8 * lk;
            const Bf16* src = k0 < k1 ? in1 + rc * k1 + k0 : in2 + rc * k2 + (k0 - k1);
            const bf16x8_t a = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(src));
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8_t bv =
                    __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&Wt[(16 * t + li) * kLd + k0]));
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t node = base + wave * 16 + 4 * lk + r;
                if (node < N) stv(out[node * kNnH + 16 * t + li], fmaxf(acc[t][r], 0.0f));
            }
    }
}

// lane-layout probe: lane l feeds a[8l .. 8l+7] and b[8l .. 8l+7] (BF16-exact
// values) to one V_MFMA_F32_16X16X32_BF16 from zero; out[4l + r] = element r
__global__ void k_mfma_bf16_probe(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
    const int l = threadIdx.x;
    uint16_t ua[8], ub[8];
    for (int e = 0; e < 8; ++e) { ua[e] = bf16_bits(a[8 * l + e]); ub[e] = bf16_bits(b[8 * l + e]); }
    const bf16x8_t va = __builtin_bit_cast(bf16x8_t, ua);
    const bf16x8_t vb = __builtin_bit_cast(bf16x8_t, ub);
    const f32x4_t d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = d[r];
}

// out[node] = sum_{p in adj(node)} val[p] src[idx[p]] / max(deg, 1)
template <typename AT>
__global__ __launch_bounds__(256) void k_mean(int32_t N, const int64_t* __restrict__ ptr,
                                              const int32_t* __restrict__ idx, const double* __restrict__ val,
                                              const AT* __restrict__ src, AT* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int32_t node = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= N) return;
    const int64_t b = ptr[node], e = ptr[node + 1];
    float acc = 0.0f;
    for (int64_t p = b; p < e; ++p) acc += (float)val[p] * ldv(src[(int64_t)idx[p] * kNnH + lane]);
    const float deg = (float)(e - b > 1 ? e - b : 1);
    stv(out[(int64_t)node * kNnH + lane], acc / deg);
}

template <typename AT>
__global__ __launch_bounds__(256) void k_output(int32_t m, const AT* __restrict__ z, const float* __restrict__ wo,
                                                const float* __restrict__ bo, const double* __restrict__ g,
                                                double rs, float* __restrict__ y, double* __restrict__ x0) {
    const int lane = threadIdx.x & 63;
    const int32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= m) return;
    const float v = wave_sum_f(ldv(z[(int64_t)row * kNnH + lane]) * wo[lane]);
    if (lane == 0) {
        const float yy = v + bo[0];
        y[row] = yy;
        if (x0) x0[row] = g[row] > 0 ? (double)yy * rs / g[row] : 0.0;
    }
}

// ----------------------------------------------------------------- loss
__global__ __launch_bounds__(256) void k_loss(int32_t m, const float* __restrict__ y, const double* __restrict__ dy,
                                              const double* __restrict__ g, double rs, double* __restrict__ part) {
    __shared__ double s0[kB], s1[kB];
    double a = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < m; i += (int64_t)gridDim.x * kB) {
        const double t = dy[i] * g[i] / rs;
        const double e = (double)y[i] - t;
        a += e * e;
        c += t * t;
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s0[threadIdx.x] += s0[threadIdx.x + w]; s1[threadIdx.x] += s1[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = s0[0]; part[gridDim.x + blockIdx.x] = s1[0]; }
}

__global__ void k_out_grad(int32_t m, const float* __restrict__ y, const double* __restrict__ dy,
                           const double* __restrict__ g, double rs, double T, float* __restrict__ dyv) {
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const double t = dy[i] * g[i] / rs;
    dyv[i] = (float)(2.0 * ((double)y[i] - t) / T);
}

// ------------------------------------------------------------- backward
// dpre[i][o] = dyv_i wo[o] (z[i][o] > 0)
template <typename AT>
__global__ void k_out_bwd(int32_t m, const AT* __restrict__ z, const float* __restrict__ dyv,
                          const float* __restrict__ wo, float* __restrict__ dpre) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= (int64_t)m * kNnH) return;
    const int64_t i = t / kNnH;
    const int o = (int)(t - i * kNnH);
    dpre[t] = ldv(z[t]) > 0.0f ? dyv[i] * wo[o] : 0.0f;
}

// partial[blk][o] = sum_i dyv_i z[i][o] (o < 64), partial[blk][64] = sum_i dyv_i
template <typename AT>
__global__ __launch_bounds__(256) void k_out_w(int32_t m, const AT* __restrict__ z, const float* __restrict__ dyv,
                                               float* __restrict__ partial) {
    const int t = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.x * kNnChunk, b1 = std::min<int64_t>(b0 + kNnChunk, m);
    float acc = 0.0f;
    if (t < kNnH) {
        for (int64_t i = b0; i < b1; ++i) acc += dyv[i] * ldv(z[i * kNnH + t]);
    } else if (t == kNnH) {
        for (int64_t i = b0; i < b1; ++i) acc += dyv[i];
    }
    if (t <= kNnH) partial[(int64_t)blockIdx.x * (kNnH + 1) + t] = acc;
}

template <typename AT>
__global__ void k_relu_bwd(int64_t N, const AT* __restrict__ h, const float* __restrict__ dh, float* __restrict__ dpre) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t >= N * kNnH) return;
    dpre[t] = ldv(h[t]) > 0.0f ? dh[t] : 0.0f;
}

// din[node][k] = sum_o dpre[node][o] W[k][o], k < k1 -> din1 (+= if acc1), else din2
__global__ __launch_bounds__(256) void k_dense_bwd_in(int64_t N, const float* __restrict__ dpre,
                                                      const float* __restrict__ W, int k1, int k2,
                                                      float* __restrict__ din1, float* __restrict__ din2, int acc1) {
    __shared__ float Ws[2 * kNnH * kNnH];
    const int K = k1 + k2;
    for (int t = threadIdx.x; t < K * kNnH; t += kB) Ws[t] = W[t];
    __syncthreads();
    const int64_t total = N * K;
    for (int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x; t < total; t += (int64_t)gridDim.x * kB) {
        const int64_t node = t / K;
        const int k = (int)(t - node * K);
        const float* dp = dpre + node * kNnH;
        float acc = 0.0f;
        for (int o = 0; o < kNnH; ++o) acc += dp[o] * Ws[k * kNnH + o];
        if (k < k1) {
            float* d = din1 + node * k1 + k;
            *d = acc1 ? *d + acc : acc;
        } else {
            din2[node * k2 + (k - k1)] = acc;
        }
    }
}

// partial[blk][k][o] = sum over the block's nodes of x_k dpre_o, x = [in1 in2 1]
template <typename AT, typename IT>
__global__ __launch_bounds__(256) void k_dense_bwd_w(int64_t N, const IT* __restrict__ in1, int k1,
                                                     const AT* __restrict__ in2, int k2,
                                                     const float* __restrict__ dpre, float* __restrict__ partial) {
    constexpr int kStage = 8;
    constexpr int kMaxU = (2 * kNnH + 1) * kNnH / kB + 1;   // 33
    __shared__ float xs[kStage][2 * kNnH + 1];
    __shared__ float ds[kStage][kNnH];
    const int K = k1 + k2, KP = K + 1;
    const int npair = KP * kNnH;
    const int t = threadIdx.x;
    float acc[kMaxU];
#pragma unroll
    for (int u = 0; u < kMaxU; ++u) acc[u] = 0.0f;
    const int64_t b0 = (int64_t)blockIdx.x * kNnChunk, b1 = std::min<int64_t>(b0 + kNnChunk, N);
    for (int64_t base = b0; base < b1; base += kStage) {
        for (int s = t; s < kStage * KP; s += kB) {
            const int q = s / KP, k = s - q * KP;
            const int64_t node = base + q;
            float v = 0.0f;
            if (node < b1) {
                if (k < k1) v = ldv(in1[node * k1 + k]);
                else if (k < K) v = ldv(in2[node * k2 + (k - k1)]);
                else v = 1.0f;
            }
            xs[q][k] = v;
        }
        for (int s = t; s < kStage * kNnH; s += kB) {
            const int q = s / kNnH, o = s - q * kNnH;
            const int64_t node = base + q;
            ds[q][o] = node < b1 ? dpre[node * kNnH + o] : 0.0f;
        }
        __syncthreads();
        for (int q = 0; q < kStage; ++q) {
#pragma unroll
            for (int u = 0; u < kMaxU; ++u) {
                const int p = t + kB * u;
                if (p < npair) acc[u] += xs[q][p >> 6] * ds[q][p & 63];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kMaxU; ++u) {
        const int p = t + kB * u;
        if (p < npair) partial[(int64_t)blockIdx.x * npair + p] = acc[u];
    }
}

__global__ void k_reduce_blocks(int npar, int nblk, const float* __restrict__ partial, float* __restrict__ out) {
    const int p = blockIdx.x * kB + threadIdx.x;
    if (p >= npar) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * npar + p];
    out[p] = s;
}

// din[src][f] += sum_{p in adjT(src)} valT[p] dout[idxT[p]][f] / deg(idxT[p])
__global__ __launch_bounds__(256) void k_mean_bwd(int32_t Nsrc, const int64_t* __restrict__ ptrT,
                                                  const int32_t* __restrict__ idxT, const double* __restrict__ valT,
                                                  const int64_t* __restrict__ ptrF, const float* __restrict__ dout,
                                                  float* __restrict__ din) {
    const int lane = threadIdx.x & 63;
    const int32_t node = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (node >= Nsrc) return;
    float acc = 0.0f;
    for (int64_t p = ptrT[node]; p < ptrT[node + 1]; ++p) {
        const int32_t j = idxT[p];
        const int64_t deg = ptrF[j + 1] - ptrF[j];
        acc += (float)valT[p] * dout[(int64_t)j * kNnH + lane] / (float)(deg > 1 ? deg : 1);
    }
    din[(int64_t)node * kNnH + lane] += acc;
}

__global__ void k_adam(int npar, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ mo,
                       float* __restrict__ ve, float lr, float b1, float b2, float eps, float c1, float c2,
                       int bf16, float* __restrict__ pw) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= npar) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float np = p[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    p[i] = np;
    pw[i] = bf16 ? round_bf16(np) : np;
}

__global__ void k_finite(int32_t m, const double* __restrict__ x, int32_t* __restrict__ flag) {
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i < m && !isfinite(x[i])) *flag = 1;
}

}  // namespace

void nn_features(int32_t m, int32_t n, const double* g, const double* r, const double* dinv, double rs, double mu,
                 double nu, float* xr, float* xc, hipStream_t st) {
    help::launch_kernel(k_feat_row, grid_for(m), kB, 0, st, m, g, r, rs, mu, xr);
    help::launch_kernel(k_feat_col, grid_for(n), kB, 0, st, n, dinv, nu, xc);
}

void nn_logsum_partials(int64_t n, const double* x, int mode, double* part, int nblk, hipStream_t st) {
    help::launch_kernel(k_logsum, nblk, kB, 0, st, n, x, mode, part);
}

template <typename AT, typename IT>
void nn_dense(int64_t N, const IT* in1, int k1, const AT* in2, int k2, const float* W, const float* b, AT* out,
              hipStream_t st) {
    if (N <= 0) return;
    if constexpr (std::is_same<AT, Bf16>::value && std::is_same<IT, Bf16>::value) {
        if ((k1 + k2) % 32 == 0 && k1 % 8 == 0 && k1 + k2 <= 2 * kNnH) {
            help::launch_kernel(k_dense_mfma, grid_cap((N + 63) / 64, 2048), kB, 0, st, N, in1, k1, in2, k2, W, b, out);
            return;
        }
    }
    help::launch_kernel(k_dense<AT, IT>, grid_cap((N + 3) / 4, 2048), kB, 0, st, N, in1, k1, in2, k2, W, b, out);
}

void nn_mfma_bf16_probe(const float* a, const float* b, float* out, hipStream_t st) {
    help::launch_kernel(k_mfma_bf16_probe, 1, 64, 0, st, a, b, out);
}

template <typename AT>
void nn_mean(int32_t N, const int64_t* ptr, const int32_t* idx, const double* val, const AT* src, AT* out,
             hipStream_t st) {
    if (N > 0) help::launch_kernel(k_mean<AT>, (N + 3) / 4, kB, 0, st, N, ptr, idx, val, src, out);
}

template <typename AT>
void nn_output(int32_t m, const AT* z, const float* wo, const float* bo, const double* g, double rs, float* y,
               double* x0, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_output<AT>, (m + 3) / 4, kB, 0, st, m, z, wo, bo, g, rs, y, x0);
}

void nn_loss_partials(int32_t m, const float* y, const double* dy, const double* g, double rs, double* part,
                      int nblk, hipStream_t st) {
    help::launch_kernel(k_loss, nblk, kB, 0, st, m, y, dy, g, rs, part);
}

void nn_output_grad(int32_t m, const float* y, const double* dy, const double* g, double rs, double T, float* dyv,
                    hipStream_t st) {
    if (m > 0) help::launch_kernel(k_out_grad, grid_for(m), kB, 0, st, m, y, dy, g, rs, T, dyv);
}

template <typename AT>
void nn_out_bwd(int32_t m, const AT* z, const float* dyv, const float* wo, float* dpre, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_out_bwd<AT>, grid_for((int64_t)m * kNnH), kB, 0, st, m, z, dyv, wo, dpre);
}

template <typename AT>
void nn_out_w(int32_t m, const AT* z, const float* dyv, float* partial, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_out_w<AT>, (m + kNnChunk - 1) / kNnChunk, kB, 0, st, m, z, dyv, partial);
}

template <typename AT>
void nn_relu_bwd(int64_t N, const AT* h, const float* dh, float* dpre, hipStream_t st) {
    if (N > 0) help::launch_kernel(k_relu_bwd<AT>, grid_for(N * kNnH), kB, 0, st, N, h, dh, dpre);
}

void nn_dense_bwd_in(int64_t N, const float* dpre, const float* W, int k1, int k2, float* din1, float* din2,
                     bool accumulate1, hipStream_t st) {
    if (N > 0)
        help::launch_kernel(k_dense_bwd_in, grid_cap((N * (k1 + k2) + kB - 1) / kB, 4096), kB, 0, st, N, dpre, W, k1, k2, din1, din2,
                                                                                    accumulate1 ? 1 : 0);
}

template <typename AT, typename IT>
void nn_dense_bwd_w(int64_t N, const IT* in1, int k1, const AT* in2, int k2, const float* dpre, float* partial,
                    hipStream_t st) {
    if (N > 0)
        help::launch_kernel(k_dense_bwd_w<AT, IT>, (int)((N + kNnChunk - 1) / kNnChunk), kB, 0, st, N, in1, k1, in2, k2, dpre,
                                                                                   partial);
}

void nn_reduce_blocks(int npar, int nblk, const float* partial, float* out, hipStream_t st) {
    help::launch_kernel(k_reduce_blocks, grid_for(npar), kB, 0, st, npar, nblk, partial, out);
}

void nn_mean_bwd(int32_t Nsrc, const int64_t* ptrT, const int32_t* idxT, const double* valT, const int64_t* ptrF,
                 const float* dout, float* din, hipStream_t st) {
    if (Nsrc > 0) help::launch_kernel(k_mean_bwd, (Nsrc + 3) / 4, kB, 0, st, Nsrc, ptrT, idxT, valT, ptrF, dout, din);
}

void nn_adam(int npar, float* p, const float* g, float* mo, float* ve, float lr, float b1, float b2, float eps,
             int t, int bf16, float* pw, hipStream_t st) {
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    help::launch_kernel(k_adam, grid_for(npar), kB, 0, st, npar, p, g, mo, ve, lr, b1, b2, eps, c1, c2, bf16, pw);
}

void nn_check_finite(int32_t m, const double* x, int32_t* flag, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_finite, grid_for(m), kB, 0, st, m, x, flag);
}

#define NN_INST(AT)                                                                                              \
    template void nn_dense<AT, AT>(int64_t, const AT*, int, const AT*, int, const float*, const float*, AT*,       \
                                   hipStream_t);                                                                 \
    template void nn_mean<AT>(int32_t, const int64_t*, const int32_t*, const double*, const AT*, AT*, hipStream_t); \
    template void nn_output<AT>(int32_t, const AT*, const float*, const float*, const double*, double, float*,     \
                                double*, hipStream_t);                                                           \
    template void nn_out_bwd<AT>(int32_t, const AT*, const float*, const float*, float*, hipStream_t);            \
    template void nn_out_w<AT>(int32_t, const AT*, const float*, float*, hipStream_t);                           \
    template void nn_relu_bwd<AT>(int64_t, const AT*, const float*, float*, hipStream_t);                        \
    template void nn_dense_bwd_w<AT, AT>(int64_t, const AT*, int, const AT*, int, const float*, float*, hipStream_t);
NN_INST(float)
NN_INST(Bf16)
template void nn_dense<Bf16, float>(int64_t, const float*, int, const Bf16*, int, const float*, const float*, Bf16*,
                                    hipStream_t);
template void nn_dense_bwd_w<Bf16, float>(int64_t, const float*, int, const Bf16*, int, const float*, float*,
                                          hipStream_t);
#undef NN_INST

}  // namespace help
