// Bipartite message-passing network that predicts the PCG initial guess
// dy ~ (A D^-1 A^T)^-1 r from the row/column graph of A (private to the
// backend).  Layout and math:
//
//   features  row i: [r_i/rs, (log g_i - mu)/10, clamp(g0 r_i/(rs g_i), +-10)/10, 1]
//             col j: [(log dinv_j - nu)/10, 1]
//             g = diag(A D^-1 A^T), rs = rms(r), mu = mean log g, nu = mean log dinv,
//             g0 = exp(mu)
//   hr0 = relu(xr Wr0 + br0), hc0 = relu(xc Wc0 + bc0)                  (width 64)
//   L1: ac = meanT(hr0);  hc1 = relu([hc0 ac] W1 + b1)    (column update)
//   L2: ar = mean(hc1);   hr1 = relu([hr0 ar] W2 + b2)    (row update)
//   L3: ac = meanT(hr1);  hc2 = relu([hc1 ac] W3 + b3)
//   L4: ar = mean(hc2);   hr2 = relu([hr1 ar] W4 + b4)
//   z = relu(hr2 Wh + bh), y_i = z_i . wo + bo,  x0_i = y_i rs / g_i
//   mean(h)_i = sum_j A_ij h_j / deg_i over row i; meanT over column j.
//   Loss: sum_i (y_i - t_i)^2 / sum_i t_i^2 with t_i = dy_i g_i / rs.
//
// Activations are stored in BF16 (or FP32 for gradient checks); every
// product accumulates in FP32; the weights used by the kernels are the FP32
// master weights rounded to BF16 in the BF16 mode, where the message layers
// and the head (K = 128 and 64) run on BF16 MFMA.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace help {

constexpr int kNnH = 64;       // hidden width
constexpr int kNnFR = 4;       // row input features
constexpr int kNnFC = 2;       // column input features
constexpr int kNnChunk = 1024; // nodes per block of the weight-gradient partials

// offsets of the parameter blocks in the flat FP32 parameter vector
struct NnLayout {
    static constexpr int Wr0 = 0, br0 = Wr0 + kNnFR * kNnH;
    static constexpr int Wc0 = br0 + kNnH, bc0 = Wc0 + kNnFC * kNnH;
    static constexpr int W1 = bc0 + kNnH;                       // 4 message layers
    static constexpr int LayerSize = 2 * kNnH * kNnH + kNnH;    // W (128x64) + b
    static constexpr int Wh = W1 + 4 * LayerSize, bh = Wh + kNnH * kNnH;
    static constexpr int wo = bh + kNnH, bo = wo + kNnH;
    static constexpr int total = bo + 1;
    static constexpr int W(int l) { return W1 + (l - 1) * LayerSize; }
    static constexpr int b(int l) { return W(l) + 2 * kNnH * kNnH; }
};

struct Bf16 { uint16_t v; };

// ---- kernels (kernels_mpnn.hip); AT = float or Bf16 activation storage ----
void nn_features(int32_t m, int32_t n, const double* g, const double* r, const double* dinv, double rs,
                 double mu, double nu, float* xr, float* xc, hipStream_t st);
void nn_logsum_partials(int64_t n, const double* x, int mode, double* part, int nblk, hipStream_t st);
template <typename AT, typename IT>
void nn_dense(int64_t N, const IT* in1, int k1, const AT* in2, int k2, const float* W, const float* b, AT* out,
              hipStream_t st);
// one V_MFMA_F32_16X16X32_BF16 from zero: lane l feeds a[8l..8l+7], b[8l..8l+7]; out[4l+r]
void nn_mfma_bf16_probe(const float* a, const float* b, float* out, hipStream_t st);
template <typename AT>
void nn_mean(int32_t N, const int64_t* ptr, const int32_t* idx, const double* val, const AT* src, AT* out,
             hipStream_t st);
template <typename AT>
void nn_output(int32_t m, const AT* z, const float* wo, const float* bo, const double* g, double rs, float* y,
               double* x0, hipStream_t st);
void nn_loss_partials(int32_t m, const float* y, const double* dy, const double* g, double rs, double* part,
                      int nblk, hipStream_t st);
void nn_output_grad(int32_t m, const float* y, const double* dy, const double* g, double rs, double T,
                    float* dyv, hipStream_t st);
template <typename AT>
void nn_out_bwd(int32_t m, const AT* z, const float* dyv, const float* wo, float* dpre, hipStream_t st);
// partial[blk][0..63] = sum dyv_i z_i, partial[blk][64] = sum dyv_i (blocks of kNnChunk rows)
template <typename AT>
void nn_out_w(int32_t m, const AT* z, const float* dyv, float* partial, hipStream_t st);
template <typename AT>
void nn_relu_bwd(int64_t N, const AT* h, const float* dh, float* dpre, hipStream_t st);
// dIn = dpre W^T (columns split into the two inputs); accumulate adds into dIn
void nn_dense_bwd_in(int64_t N, const float* dpre, const float* W, int k1, int k2, float* din1, float* din2,
                     bool accumulate1, hipStream_t st);
// partial (per block of kNnChunk nodes) of [in1 in2 1]^T dpre, (k1+k2+1) x 64
template <typename AT, typename IT>
void nn_dense_bwd_w(int64_t N, const IT* in1, int k1, const AT* in2, int k2, const float* dpre, float* partial,
                    hipStream_t st);
void nn_reduce_blocks(int npar, int nblk, const float* partial, float* out, hipStream_t st);
// din[src] += sum over the transposed adjacency of val * dout[dst] / deg(dst)
void nn_mean_bwd(int32_t Nsrc, const int64_t* ptrT, const int32_t* idxT, const double* valT,
                 const int64_t* ptrF, const float* dout, float* din, hipStream_t st);
void nn_adam(int npar, float* p, const float* g, float* mo, float* ve, float lr, float b1, float b2, float eps,
             int t, int bf16, float* pw, hipStream_t st);
void nn_check_finite(int32_t m, const double* x, int32_t* flag, hipStream_t st);

}  // namespace help
