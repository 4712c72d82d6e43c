// Host side of the matrix-free PCG mode (HELP UsePCG_CPU.cpp:91-178 and
// PCG.cpp:38-395): the partial-Cholesky preconditioner is built once per
// renew (d is the same for the three solves of an iteration; the reference
// rebuilds it inside every solve, PCG.cpp:301), then each solve runs FP64
// conjugate gradients on A D^-1 A^T (gamma = delta = 0, UsePCG_CPU.cpp:154-155)
// with the preconditioner applied from the FP32 copy of the factor.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "handle.h"
#include "pcg.h"

namespace {
// Scratch of the allocation-free preconditioner apply
// (help_gpu_pcg_apply_noalloc): the apply's y1/w vectors (2 * kPcgMaxK) and
// its per-block partial sums (K * ceil(m / kLtRows)).  Every GPU holds one
// copy until the process ends; calls that overlap on one GPU share it and
// race.  The kernels write every entry they read, so nothing needs clearing.
constexpr int64_t kNoAllocScratch = int64_t(1) << 20;   // doubles, 8 MiB
__device__ __align__(256) double scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per
// device (the host-side name of the array is not a device address)
double* scratch_ptr() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw help::HipError("device index out of range", HELP_ERR_HIP);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double*)p;
    }
    return g_scratch_dev[dev];
}
}  // namespace

namespace help {

struct PcgState {
    int32_t K = 0;
    DBuf<double> cg, part;           // CG scalars (kCgSlots) and reduction partials
    double* hpin = nullptr;          // pinned: the scalars of the last two iterations
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PcgState() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (hpin) (void)hipHostFree(hpin);
    }
    DBuf<double> L, D, S, w, g, sc, L11, X, y1, ltpart, pval;
    DBuf<float> L32;
    DBuf<int32_t> perm, pos, col, pidx;
    DBuf<double> x, r, z, p, q, t;
    std::vector<int32_t> iota;
};

void pcg_free(PcgState* p) { delete p; }

void pcg_setup(help_gpu* h) {
    auto* P = new PcgState();
    h->pcg = P;
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    P->K = std::max(0, std::min<int32_t>(std::min(h->opts.pcg_k, kPcgMaxK), m));
    const int64_t K = std::max<int32_t>(P->K, 1);
    P->L.alloc((size_t)m * K);
    if (h->opts.pcg_fp32_precond) P->L32.alloc((size_t)m * K);
    P->D.alloc(K);
    P->w.alloc(K);
    P->S.alloc(m);
    P->g.alloc(m);
    P->sc.alloc(n);
    P->L11.alloc((size_t)K * K);
    P->X.alloc((size_t)K * K);
    P->y1.alloc(2 * kPcgMaxK);
    P->ltpart.alloc((size_t)K * ((m + kLtRows - 1) / kLtRows));
    P->pval.alloc(kArgmaxBlocks);
    P->pidx.alloc(kArgmaxBlocks);
    P->perm.alloc(m);
    P->pos.alloc(m);
    P->col.alloc(1);
    P->x.alloc(m); P->r.alloc(m); P->z.alloc(m); P->p.alloc(m); P->q.alloc(m);
    P->t.alloc(n);
    P->cg.alloc(kCgSlots);
    P->part.alloc(kDotPartials);
    HIPCHK(hipHostMalloc((void**)&P->hpin, 2 * kCgSlots * sizeof(double), hipHostMallocDefault));
    for (auto& e : P->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    P->iota.resize(m);
    for (int32_t i = 0; i < m; ++i) P->iota[i] = i;
    HIPCHK(hipMemsetAsync(P->sc.p, 0, sizeof(double) * n, st));
    HIPCHK(hipStreamSynchronize(st));
}

// Greedy partial LDL^T of A D^-1 A^T with K pivots (PCG.cpp:38-103), the
// inverse of its unit-lower leading block (PCG.cpp:119-133), and the FP32
// copy of the factor for the apply.  d_dinv must hold 1/d.
void pcg_renew(help_gpu* h) {
    PcgState* P = h->pcg;
    hipStream_t st = h->st;
    const int32_t m = h->m;
    const int32_t K = P->K;
    // Schur diagonal: diag(A D^-1 A^T) (PCG.cpp:40-46)
    launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_dinv.p, P->S.p, true, st);
    HIPCHK(hipMemcpyAsync(P->perm.p, P->iota.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(P->pos.p, P->iota.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    for (int32_t k = 0; k < K; ++k) {
        launch_pc_argmax(m, k, P->S.p, P->pval.p, P->pidx.p, P->perm.p, P->pos.p, P->col.p, st);
        // column of A D^-1 A^T at the pivot row: A (D^-1 a_col^T) (PCG.cpp:63-79)
        launch_pc_scatter(P->col.p, h->max_row_nnz, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_dinv.p,
                          P->sc.p, false, st);
        launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, P->sc.p, P->g.p, false, st);
        launch_pc_scatter(P->col.p, h->max_row_nnz, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_dinv.p,
                          P->sc.p, true, st);
        // left-looking update against the previous columns (PCG.cpp:85-102)
        launch_pc_pivot(m, k, P->col.p, P->g.p, P->L.p, 0.0, P->D.p, P->w.p, st);
        launch_pc_column(m, k, P->col.p, P->pos.p, P->g.p, P->w.p, P->D.p, P->L.p, P->S.p, st);
    }
    launch_pc_l11(m, K, P->perm.p, P->L.p, P->L11.p, P->X.p, st);
    if (h->opts.pcg_fp32_precond) launch_to_f32((int64_t)m * K, P->L.p, P->L32.p, st);
    HIPCHK(hipGetLastError());
}

// per pivot k: the product with A (12 nnz + 8 (n + m)), the left-looking
// update reading k previous columns (8 m k) and writing one (8 m); the FP32
// copy at the end (12 m K)
double pcg_build_bytes(const help_gpu* h) {
    const double nnz = (double)h->nnz, m = h->m, n = h->n, K = h->pcg ? h->pcg->K : 0;
    return K * (12.0 * nnz + 8.0 * (n + m) + 8.0 * m) + 8.0 * m * K * (K - 1) / 2 + 12.0 * m * K;
}

// per iteration: A^T p scaled (12 nnz + 8 m + 16 n), A t (12 nnz + 8 n + 8 m),
// the preconditioner (two passes over the m x K factor: 8 m K in FP32,
// 16 m K in FP64) and about ten m-vector streams
double pcg_iteration_bytes(const help_gpu* h) {
    const double nnz = (double)h->nnz, m = h->m, n = h->n, K = h->pcg ? h->pcg->K : 0;
    const double w = h->opts.pcg_fp32_precond ? 4.0 : 8.0;
    return 24.0 * nnz + 24.0 * n + 16.0 * m + 2.0 * w * m * K + 80.0 * m;
}

static double dot(help_gpu* h, const double* a, const double* b, int64_t n) {
    launch_dot(a, b, n, h->d_part.p, h->d_scal.p, h->st);
    double v = 0;
    HIPCHK(hipMemcpyAsync(&v, h->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return v;
}

// Conjugate gradients (PCG.cpp:296-395) on A D^-1 A^T x = rhs from x0 (or
// zero), all vectors on the device in the original row order.  Stops at
// ||r|| <= pcg_tol ||r0|| or pcg_max_iter iterations; x2 <- x.  Returns the
// iteration count; rel is the final ||r|| / ||r0||.
//
// The scalars of the iteration (p.q, r.r, r.z, alpha, beta) stay on the
// device (launch_cg_step1/2): no host round trip inside an iteration.  The
// host copies the state of every iteration to pinned memory and reads it one
// iteration later, while the next one is already queued, and stops queueing
// once it sees convergence or breakdown; the one extra iteration in flight
// then changes nothing (its kernels see the stopped state and return).
int pcg_solve(help_gpu* h, const double* rhs, const double* x0, double* x2, double& rel) {
    PcgState* P = h->pcg;
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    double *x = P->x.p, *r = P->r.p, *z = P->z.p, *p = P->p.p, *q = P->q.p, *t = P->t.p;
    auto matvec = [&](const double* v, double* out) {   // out = A D^-1 A^T v
        launch_csc_t_scaled(n, h->d_colptr.p, h->d_rowind.p, h->d_val.p, h->d_dinv.p, v, t, st);
        launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, t, out, false, st);
    };
    auto precond = [&](const double* v, double* out) { pcg_precond(h, v, out); };
    const double bb = dot(h, rhs, rhs, m);
    if (x0) {
        HIPCHK(hipMemcpyAsync(x, x0, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        matvec(x, q);
        launch_sub(m, rhs, q, r, st);
    } else {
        HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * m, st));
        HIPCHK(hipMemcpyAsync(r, rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    }
    double rr = dot(h, r, r, m);
    if (x0 && !(rr <= bb)) {   // a guess worse than zero (or not finite): start from zero
        HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * m, st));
        HIPCHK(hipMemcpyAsync(r, rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        rr = bb;
    }
    // the reference's stop is relative to ||r0|| (PCG.cpp:328); with a
    // warm start the relative target is taken against ||rhs|| so that a
    // good guess cannot loosen it
    const double r0 = std::sqrt(x0 ? std::max(bb, 0.0) : rr);
    const double target = h->opts.pcg_tol * r0;
    int it = 0;
    if (r0 == 0.0) {
        rel = 0;
    } else if (!(std::sqrt(rr) > target) || h->opts.pcg_max_iter <= 0) {
        rel = std::sqrt(rr) / r0;
    } else {
        precond(r, z);
        HIPCHK(hipMemcpyAsync(p, z, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        const double rz = dot(h, r, z, m);
        double* sc = P->cg.p;
        launch_cg_init(sc, rr, rz, st);
        const int maxit = h->opts.pcg_max_iter;
        int queued = 0;
        for (; queued < maxit; ++queued) {
            const int cur = kCgRZ + (queued & 1), next = kCgRZ + ((queued + 1) & 1);
            matvec(p, q);
            launch_cg_step1(m, sc, cur, target, p, q, x, r, P->part.p, st);
            precond(r, z);
            launch_cg_step2(m, sc, cur, next, r, z, p, P->part.p, st);
            double* hp = P->hpin + kCgSlots * (queued & 1);
            HIPCHK(hipMemcpyAsync(hp, sc, sizeof(double) * kCgSlots, hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(P->ev[queued & 1], st));
            if (queued >= 1) {   // the state after the previous iteration
                HIPCHK(hipEventSynchronize(P->ev[(queued - 1) & 1]));
                if (P->hpin[kCgSlots * ((queued - 1) & 1) + kCgState] != 0.0) { ++queued; break; }
            }
        }
        HIPCHK(hipStreamSynchronize(st));
        const double* fin = P->hpin + kCgSlots * ((queued - 1) & 1);
        it = (int)fin[kCgIters];
        rel = std::sqrt(fin[kCgRR]) / r0;
    }
    HIPCHK(hipMemcpyAsync(x2, x, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipGetLastError());
    return it;
}

// z = M^-1 v with the preconditioner of the last renew (device vectors,
// original row order), as pcg_solve applies it
void pcg_precond(help_gpu* h, const double* v, double* out) {
    PcgState* P = h->pcg;
    if (h->opts.pcg_fp32_precond)
        launch_pc_apply<float>(h->m, P->K, P->perm.p, P->pos.p, P->L32.p, P->X.p, P->S.p, v, P->y1.p,
                               P->ltpart.p, out, h->st);
    else
        launch_pc_apply<double>(h->m, P->K, P->perm.p, P->pos.p, P->L.p, P->X.p, P->S.p, v, P->y1.p,
                                P->ltpart.p, out, h->st);
}

// pcg_precond with its scratch in the file-scope `scratch` array: no
// allocation, device vectors only
void pcg_precond_noalloc(help_gpu* h, const double* v, double* out) {
    PcgState* P = h->pcg;
    const int64_t nb = ((int64_t)h->m + kLtRows - 1) / kLtRows;
    const int64_t need = 2 * (int64_t)kPcgMaxK + (int64_t)std::max(P->K, 1) * nb;
    if (need > kNoAllocScratch)
        throw std::runtime_error("pcg_apply_noalloc: needs " + std::to_string(need) +
                                 " scratch doubles (2 * 1024 + K * ceil(m / 2048)), above the maximum " +
                                 std::to_string(kNoAllocScratch));
    double* s = scratch_ptr();
    if (h->opts.pcg_fp32_precond)
        launch_pc_apply<float>(h->m, P->K, P->perm.p, P->pos.p, P->L32.p, P->X.p, P->S.p, v, s,
                               s + 2 * kPcgMaxK, out, h->st);
    else
        launch_pc_apply<double>(h->m, P->K, P->perm.p, P->pos.p, P->L.p, P->X.p, P->S.p, v, s,
                                s + 2 * kPcgMaxK, out, h->st);
}

void pcg_download(help_gpu* h, int32_t* K, int32_t* perm, double* D, double* S, double* L) {
    PcgState* P = h->pcg;
    hipStream_t st = h->st;
    const int32_t m = h->m;
    if (K) *K = P->K;
    if (perm) HIPCHK(hipMemcpyAsync(perm, P->perm.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, st));
    if (D && P->K) HIPCHK(hipMemcpyAsync(D, P->D.p, sizeof(double) * P->K, hipMemcpyDeviceToHost, st));
    if (S) HIPCHK(hipMemcpyAsync(S, P->S.p, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    if (L && P->K) HIPCHK(hipMemcpyAsync(L, P->L.p, sizeof(double) * (size_t)m * P->K, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
}

}  // namespace help
