// Supernodal multifrontal Cholesky on CDNA4.
//
// Storage per supernode s (ncol n, nrow c, structure rows sorted):
//   L panel  c x n, column-major (lda = c), rows = the structure of s;
//   U        (c-n) x (c-n), column-major lower, the update ("contribution")
//            matrix passed to the parent by extend-add;
//   Dinv     the inverse of every 64-column diagonal block of L11 (nb x nb,
//            column-major), used by the panel solve and the triangular solves.
// The symbolic schedule processes the elimination tree level by level;
// within a level every front is independent.  A parent pulls its children's
// update matrices one child after the other, so the summation order is fixed
// and no atomics touch the factor.
//
// Small fronts (c <= 256, c*n <= 7680): one workgroup per front does the
// extend-add, the dense partial factorization in LDS, the diagonal-block
// inverses and the update SYRK.
// Large fronts: extend-add by column chunks, then for each 64-column panel a
// two-wave diagonal-block Cholesky + inverse, the panel rows times the
// inverse on MFMA, an MFMA GEMM on the trailing panel columns, and finally
// one MFMA SYRK for U with the full inner dimension n.
#include <type_traits>

#include "device.h"

namespace help {

namespace {

template <typename T> struct Huge;
template <> struct Huge<double> { static constexpr double v = 1e128; };
template <> struct Huge<float> { static constexpr float v = 1e30f; };

// pivot policy: a non-positive (or NaN) pivot is replaced by a huge value, so
// the column of L becomes ~0 (that row is dropped from the solve), and counted.
template <typename T>
__device__ inline T safe_sqrt_pivot(T d, int32_t* err, bool count) {
    if (!(d > T(0))) {
        if (count) atomicAdd(err, 1);
        d = Huge<T>::v;
    }
    return sqrt(d);
}

// orders the LDS accesses of one wave (no other wave takes part)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int64_t dinv_block_offset(const DevSym& sym, int32_t s, int32_t b) {
    return sym.dinvoff[s] + (int64_t)(b / kTile) * kTile * kTile;
}

// X = L^-1 of the factored lower-triangular nb x nb block at Lsrc (LDS,
// leading dimension lds) by row operations on the identity (256 threads),
// written to Db (column-major nb x nb).  X is LDS scratch of nb*(nb+1).
template <typename T>
__device__ void invert_lower_block(const T* Lsrc, int32_t lds, int32_t nb, T* __restrict__ Db, T* X) {
    const int tid = threadIdx.x;
    const int ldx = nb + 1;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        int j = idx / nb, i = idx - j * nb;
        X[j * ldx + i] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const T djj = Lsrc[j * lds + j];
        if (tid <= j) X[tid * ldx + j] /= djj;      // row j, columns k <= j
        __syncthreads();
        // rows i > j, columns k <= j:  X[i][k] -= L[i][j] X[j][k]
        const int cnt = (nb - j - 1) * (j + 1);
        for (int idx = tid; idx < cnt; idx += 256) {
            int k = idx / (nb - j - 1), i = j + 1 + idx - k * (nb - j - 1);
            X[k * ldx + i] -= Lsrc[j * lds + i] * X[k * ldx + j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * nb; idx += 256) {
        int j = idx / nb, i = idx - j * nb;
        Db[j * nb + i] = (i >= j) ? X[j * ldx + i] : T(0);
    }
}

// ------------------------------------------------------------------ small
// Dynamic LDS: the c x n panel followed by the (nb x nb+1) inverse scratch.
template <typename T>
__global__ __launch_bounds__(256) void k_small_fronts(const int32_t* __restrict__ list, DevSym sym,
                                                      T* __restrict__ L, T* __restrict__ U,
                                                      T* __restrict__ Dinv,
                                                      int32_t* __restrict__ err) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ T piv;
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    T* Lp = L + sym.loff[s];
    T* Up = U + sym.uoff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int32_t cn = c * n;
    for (int32_t idx = tid; idx < cn; idx += nt) F[idx] = Lp[idx];
    // the update matrix starts from zero (its lower part: the upper part is
    // never used)
    for (int32_t idx = tid; idx < cu * cu; idx += nt) {
        const int32_t jj = idx / cu;
        if (idx - jj * cu >= jj) Up[idx] = T(0);
    }
    __syncthreads();
    // extend-add the children, one after the other
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chn = sym.sfirst[ch + 1] - sym.sfirst[ch];
        const int32_t chc = sym.snrow[ch] - chn;
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const T* Uc = U + sym.uoff[ch];
        // four entries per thread in flight: every load (the child's entry,
        // the relative map, the parent's update-matrix entry) before the
        // adds; the relative map is injective, so the targets are distinct
        const int64_t tot = (int64_t)chc * chc;
        for (int64_t idx0 = tid; idx0 < tot; idx0 += 4 * nt) {
            bool ok[4];
            int32_t rr[4], cc[4];
            T v[4], old[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t idx = idx0 + (int64_t)q * nt;
                const int64_t id = idx < tot ? idx : 0;
                const int32_t jj = (int32_t)(id / chc), ii = (int32_t)(id - (int64_t)jj * chc);
                ok[q] = idx < tot && ii >= jj;
                rr[q] = rel[ii];
                cc[q] = rel[jj];
                v[q] = Uc[id];
            }
            if (cu > 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool up = ok[q] && cc[q] >= n;
                    old[q] = Up[up ? (int64_t)(cc[q] - n) * cu + (rr[q] - n) : 0];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!ok[q]) continue;
                if (cc[q] < n) F[cc[q] * c + rr[q]] += v[q];
                else Up[(int64_t)(cc[q] - n) * cu + (rr[q] - n)] = old[q] + v[q];
            }
        }
        __syncthreads();
    }
    // dense partial Cholesky of the panel, right-looking by columns
    for (int32_t j = 0; j < n; ++j) {
        if (tid == 0) {
            T d = safe_sqrt_pivot<T>(F[j * c + j], err, true);
            F[j * c + j] = d;
            piv = d;
        }
        __syncthreads();
        const T p = piv;
        for (int32_t i = j + 1 + tid; i < c; i += nt) F[j * c + i] /= p;
        __syncthreads();
        const int32_t rem = n - j - 1;
        const int32_t tot = rem * c;
        for (int32_t idx = tid; idx < tot; idx += nt) {
            int32_t k = j + 1 + idx / c, i = idx % c;
            if (i >= k) F[k * c + i] -= F[j * c + i] * F[j * c + k];
        }
        __syncthreads();
    }
    for (int32_t idx = tid; idx < cn; idx += nt) Lp[idx] = F[idx];
    // inverses of the diagonal blocks
    T* X = F + cn;
    for (int32_t b = 0; b < n; b += kTile) {
        const int32_t nb = min(kTile, n - b);
        invert_lower_block<T>(F + (int64_t)b * c + b, c, nb, Dinv + dinv_block_offset(sym, s, b), X);
    }
    // U -= L21 L21^T, four entries per thread with their loads first
    const int32_t tot = cu * cu;
    for (int32_t idx0 = tid; idx0 < tot; idx0 += 4 * nt) {
        bool ok[4];
        int32_t pos[4];
        T old[4], acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t idx = idx0 + q * nt;
            const int32_t id = idx < tot ? idx : 0;
            const int32_t jj = id / cu, ii = id - jj * cu;
            ok[q] = idx < tot && ii >= jj;
            pos[q] = id;
            old[q] = Up[id];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T a = 0;
            if (ok[q]) {
                const int32_t jj = pos[q] / cu, ii = pos[q] - jj * cu;
                for (int32_t j = 0; j < n; ++j) a += F[j * c + n + ii] * F[j * c + n + jj];
            }
            acc[q] = a;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ok[q]) Up[pos[q]] = old[q] - acc[q];
    }
}

// ------------------------------------------------------------ small, wave
// k_small_fronts for fronts of at most 64 rows: one wave per front (a row
// per lane), four fronts per workgroup, each in its own lds_elems slice of
// the dynamic LDS.  Every entry goes through the same operations in the same
// order as in k_small_fronts (extend-add child by child, right-looking
// columns, the inverse by row operations, U -= L21 L21^T), with wave-level
// ordering in place of the workgroup barriers, so the results are the same.
__device__ inline void wave_fence_all() {
    __threadfence_block();   // global read-modify-writes of one child before the next child's
    __builtin_amdgcn_wave_barrier();
}

template <typename T, int G>
__global__ __launch_bounds__(256) void k_small_wave(const int32_t* __restrict__ list, int32_t count,
                                                    int32_t lds_elems, DevSym sym, T* __restrict__ L,
                                                    T* __restrict__ U, T* __restrict__ Dinv,
                                                    int32_t* __restrict__ err) {
    extern __shared__ unsigned char smem_raw[];
    // G lanes per front (c <= G): 64, or 32 for two fronts per wave
    const int lane = threadIdx.x & (G - 1), grp = threadIdx.x / G;
    const int32_t task = blockIdx.x * (256 / G) + grp;
    if (task >= count) return;   // the fronts are independent: no barrier below
    T* F = reinterpret_cast<T*>(smem_raw) + (int64_t)grp * lds_elems;
    const int32_t s = list[task];
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    T* Lp = L + sym.loff[s];
    T* Up = U + sym.uoff[s];
    const int32_t cn = c * n;
    for (int32_t idx = lane; idx < cn; idx += G) F[idx] = Lp[idx];
    for (int32_t idx = lane; idx < cu * cu; idx += G) {   // update matrix from zero, lower part
        const int32_t jj = idx / cu;
        if (idx - jj * cu >= jj) Up[idx] = T(0);
    }
    wave_fence_all();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chn = sym.sfirst[ch + 1] - sym.sfirst[ch];
        const int32_t chc = sym.snrow[ch] - chn;
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const T* Uc = U + sym.uoff[ch];
        const int64_t tot = (int64_t)chc * chc;
        for (int64_t idx0 = lane; idx0 < tot; idx0 += 4 * G) {
            bool ok[4];
            int32_t rr[4], cc[4];
            T v[4], old[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t idx = idx0 + (int64_t)q * G;
                const int64_t id = idx < tot ? idx : 0;
                const int32_t jj = (int32_t)(id / chc), ii = (int32_t)(id - (int64_t)jj * chc);
                ok[q] = idx < tot && ii >= jj;
                rr[q] = rel[ii];
                cc[q] = rel[jj];
                v[q] = Uc[id];
            }
            if (cu > 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool up = ok[q] && cc[q] >= n;
                    old[q] = Up[up ? (int64_t)(cc[q] - n) * cu + (rr[q] - n) : 0];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (!ok[q]) continue;
                if (cc[q] < n) F[cc[q] * c + rr[q]] += v[q];
                else Up[(int64_t)(cc[q] - n) * cu + (rr[q] - n)] = old[q] + v[q];
            }
        }
        wave_fence_all();
    }
    // dense partial Cholesky of the panel, right-looking by columns
    for (int32_t j = 0; j < n; ++j) {
        const T p = safe_sqrt_pivot<T>(F[j * c + j], err, lane == 0);
        wave_sync();
        if (lane == 0) F[j * c + j] = p;
        if (lane > j && lane < c) F[j * c + lane] /= p;
        wave_sync();
        if (lane > j && lane < c) {
            for (int32_t k = j + 1; k < n && k <= lane; ++k)
                F[k * c + lane] -= F[j * c + lane] * F[j * c + k];
        }
        wave_sync();
    }
    for (int32_t idx = lane; idx < cn; idx += G) Lp[idx] = F[idx];
    // inverse of the (single, n <= 64) diagonal block, as invert_lower_block
    {
        const int32_t nb = n, ldx = nb + 1;
        T* X = F + cn;
        T* Db = Dinv + dinv_block_offset(sym, s, 0);
        for (int idx = lane; idx < nb * nb; idx += G) {
            int j = idx / nb, i = idx - j * nb;
            X[j * ldx + i] = (i == j) ? T(1) : T(0);
        }
        wave_sync();
        for (int j = 0; j < nb; ++j) {
            const T djj = F[j * c + j];
            if (lane <= j) X[lane * ldx + j] /= djj;
            wave_sync();
            const int cnt = (nb - j - 1) * (j + 1);
            for (int idx = lane; idx < cnt; idx += G) {
                int k = idx / (nb - j - 1), i = j + 1 + idx - k * (nb - j - 1);
                X[k * ldx + i] -= F[j * c + i] * X[k * ldx + j];
            }
            wave_sync();
        }
        for (int idx = lane; idx < nb * nb; idx += G) {
            int j = idx / nb, i = idx - j * nb;
            Db[j * nb + i] = (i >= j) ? X[j * ldx + i] : T(0);
        }
    }
    // U -= L21 L21^T, four entries per lane with their loads first
    const int32_t tot = cu * cu;
    for (int32_t idx0 = lane; idx0 < tot; idx0 += 4 * G) {
        bool ok[4];
        int32_t pos[4];
        T old[4], acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t idx = idx0 + q * G;
            const int32_t id = idx < tot ? idx : 0;
            const int32_t jj = id / cu, ii = id - jj * cu;
            ok[q] = idx < tot && ii >= jj;
            pos[q] = id;
            old[q] = Up[id];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T a = 0;
            if (ok[q]) {
                const int32_t jj = pos[q] / cu, ii = pos[q] - jj * cu;
                for (int32_t j = 0; j < n; ++j) a += F[j * c + n + ii] * F[j * c + n + jj];
            }
            acc[q] = a;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ok[q]) Up[pos[q]] = old[q] - acc[q];
    }
}

// ------------------------------------------------------------ extend-add
// Task = kEaCols columns [c0, c1) of a front.  The update-matrix columns of
// the range are first set to zero (lower part), then for each child in order
// the child columns that land in the range are spread over the 4 waves (two
// columns per wave), the lanes take the rows, kDepth rows per lane in flight
// (8 in the launches of large fronts, whose children have long columns).
// Children with no column in the range are skipped without a barrier.
template <typename T, int kDepth>
__global__ __launch_bounds__(256) void k_extend_add(const EaTask* __restrict__ tasks,
                                                    const EaRef* __restrict__ refs, DevSym sym,
                                                    T* __restrict__ L, T* __restrict__ U) {
    const EaTask tk = tasks[blockIdx.x];
    const int32_t n = tk.n;
    const int32_t c = tk.c;
    const int32_t cu = c - n;
    T* Lp = L + tk.loff;
    T* Up = U + tk.uoff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (tk.c1 > n) {   // uniform over the workgroup
        for (int32_t j = max(tk.c0, n); j < tk.c1; ++j) {
            T* col = Up + (int64_t)(j - n) * cu - n;   // indexed by front row
            for (int32_t i = j + (int32_t)threadIdx.x; i < c; i += 256) col[i] = T(0);
        }
        __syncthreads();
    }
    for (int32_t t = 0; t < tk.nref; ++t) {
        const EaRef rf = refs[tk.r0 + t];
        const int32_t chc = rf.chc;
        const int32_t* rel = sym.relmap + rf.reloff;
        const T* Uc = U + rf.uoff;
        const int32_t lo = rf.lo, hi2 = rf.hi;
        // each wave takes two child columns at a time, 4 rows of each per lane
        // in flight; all loads are issued before the read-modify-writes
        for (int32_t j0 = lo + wave; j0 < hi2; j0 += 8) {
            const int32_t jA = j0, jB = min(j0 + 4, hi2 - 1);
            const bool hasB = j0 + 4 < hi2;
            const int32_t ccA = rel[jA], ccB = rel[jB];
            const T* colA = Uc + (int64_t)jA * chc;
            const T* colB = Uc + (int64_t)jB * chc;
            T* dstA = (ccA < n) ? Lp + (int64_t)ccA * c : Up + (int64_t)(ccA - n) * cu - n;
            T* dstB = (ccB < n) ? Lp + (int64_t)ccB * c : Up + (int64_t)(ccB - n) * cu - n;
            const int32_t endA = chc, endB = hasB ? chc : jB;   // B empty when absent
            for (int32_t i0 = lane; jA + i0 < endA || jB + i0 < endB; i0 += 64 * kDepth) {
                int32_t ra[kDepth], rb[kDepth];
                T va[kDepth], vb[kDepth];
#pragma unroll
                for (int u = 0; u < kDepth; ++u) {
                    const int32_t ia = jA + i0 + 64 * u, ib = jB + i0 + 64 * u;
                    const bool oka = ia < endA, okb = ib < endB;
                    const int32_t ca = oka ? ia : jA, cb = okb ? ib : jB;
                    const int32_t r1 = rel[ca], r2 = rel[cb];
                    const T x1 = colA[ca], x2 = colB[cb];
                    ra[u] = oka ? r1 : -1;
                    rb[u] = okb ? r2 : -1;
                    va[u] = x1;
                    vb[u] = x2;
                }
                T da[kDepth], db[kDepth];
#pragma unroll
                for (int u = 0; u < kDepth; ++u) {
                    da[u] = dstA[ra[u] >= 0 ? ra[u] : ccA];
                    db[u] = dstB[rb[u] >= 0 ? rb[u] : ccB];
                }
#pragma unroll
                for (int u = 0; u < kDepth; ++u) {
                    if (ra[u] >= 0) dstA[ra[u]] = da[u] + va[u];
                    if (rb[u] >= 0) dstB[rb[u]] = db[u] + vb[u];
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------ diagonal block Cholesky + inverse
// Elimination with unit multipliers on the augmented block [A | I]: step j
// stores m_ij = a_ij / d_j in place, updates the trailing lower part
// (a_ik -= d_j m_ij m_kj) and applies the same row operation to the identity
// (x_ik -= m_ij x_jk).  At the end A holds L_u (unit) and D, X holds L_u^-1,
// so L = L_u D^1/2 and L^-1 = D^-1/2 L_u^-1.  The block is padded with the
// identity to 64, which leaves the leading nb x nb results unchanged.
template <typename T>
__device__ void factor_invert_block(const T* __restrict__ src, int64_t lds, int32_t nb,
                                    T* __restrict__ Ldst, T* __restrict__ Db, T* A, T* X,
                                    int32_t* err) {
    constexpr int ld = kTile + 1;
    const int tid = threadIdx.x;
    const int row = tid & 63, part = tid >> 6;
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        int j = idx >> 6, i = idx & 63;
        T v = (i == j) ? T(1) : T(0);
        if (i < nb && j < nb && i >= j) v = src[(int64_t)j * lds + i];
        A[j * ld + i] = v;
        X[j * ld + i] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = 0; j < kTile; ++j) {
        T piv = A[j * ld + j];
        if (!(piv > T(0))) {
            if (tid == 0 && j < nb) atomicAdd(err, 1);
            piv = Huge<T>::v;
        }
        __syncthreads();
        if (part == 0) {
            if (row > j) A[j * ld + row] /= piv;
            else if (row == j) A[j * ld + j] = piv;
        }
        __syncthreads();
        if (row > j) {
            const T mi = A[j * ld + row];
            for (int k = part; k <= row; k += 4) {
                if (k > j) A[k * ld + row] -= piv * mi * A[j * ld + k];
                else X[k * ld + row] -= mi * X[k * ld + j];
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        int j = idx >> 6, i = idx & 63;
        if (i >= nb || j >= nb) continue;
        const T sdj = sqrt(A[j * ld + j]);
        if (i >= j) Ldst[(int64_t)j * lds + i] = (i == j) ? sdj : A[j * ld + i] * sdj;
        Db[j * nb + i] = (i >= j) ? X[j * ld + i] / sqrt(A[i * ld + i]) : T(0);
    }
}

// Diagonal block of a panel: LDL^T of the nb x nb block (identity padding to
// 64) with L11 = L_unit diag(sqrt(d)) written in place and its inverse
// Dinv = diag(1/sqrt(d)) L_unit^-1 written to the block's Dinv slot.
//
// Two waves with 2D register blocking: lane (bi, bk) = (lane >> 3, lane & 7)
// holds the 8x8 block of rows 8bi.. and columns 8bk.. .
//  * Wave 0 factors.  Step j publishes column j of the Schur complement to
//    LDS (the lanes of column block j/8 write it), every lane reads its 8
//    row values and 8 column values and applies a -= (raw_r / d_j) raw_k to
//    its whole block.  That is exact on the trailing matrix (r, k > j); the
//    other entries it touches are never read again (finished columns are
//    already published, rows above j only own upper-triangle entries), so
//    the step needs no masks and no branches.  L11 is formed at the end from
//    the published columns.
//  * Wave 1 forms X = L_unit^-1 in the same layout: step j waits for wave 0
//    to publish step j (LDS flag; wave 0 never waits), takes l_rj for its
//    rows from the published column, and subtracts l_rj X[j][:] (row j of X
//    is final at step j and published by its owners).
// 8 steps are unrolled so every register index is static; the publishing
// stores are branch-free (non-owner lanes write a dump row), and the output
// stores are unconditional (elements outside nb x nb go to an LDS slot
// through the generic address space).
template <typename T>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_potrf_diag(const PanelTask* __restrict__ tasks, DevSym sym, T* __restrict__ L, T* __restrict__ Dinv,
                  int32_t* __restrict__ err) {
    constexpr int kS = kTile + 1;
    __shared__ T S[kTile * kS];            // input staging, then X for the Dinv output
    __shared__ T cbt[kTile * kTile];       // column j of the Schur complement at step j
    __shared__ T rps[kTile], pivs[kTile], sdv[kTile];
    __shared__ T xrow[kTile];
    __shared__ T dump[2][kTile * 8];
    __shared__ int flag;
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int nb = min(kTile, n - b);
    T* P = L + sym.loff[s] + (int64_t)b * c + b;
    T* Db = Dinv + dinv_block_offset(sym, s, b);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int bi = lane >> 3, bk = lane & 7;
    {
        // lower triangle, identity outside nb, zero above; clamped loads
        const int r = lane, rr = min(r, nb - 1);
        T v[kTile / 2];
#pragma unroll
        for (int t = 0; t < kTile / 2; ++t) {
            const int k = w + 2 * t;
            v[t] = P[(int64_t)min(k, nb - 1) * c + rr];
        }
#pragma unroll
        for (int t = 0; t < kTile / 2; ++t) {
            const int k = w + 2 * t;
            S[k * kS + r] = (r < nb && k < nb) ? ((r >= k) ? v[t] : T(0)) : ((r == k) ? T(1) : T(0));
        }
        if (tid == 0) flag = 0;
    }
    __syncthreads();
    if (w == 0) {
        T a[8][8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < 8; ++v) a[u][v] = S[(8 * bk + v) * kS + 8 * bi + u];
        int nbad = 0;
        for (int jb = 0; jb < 8; ++jb) {
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int j = 8 * jb + jj;
                T* col = cbt + j * kTile;
                T* dst = (bk == jb) ? col + 8 * bi : dump[0] + 8 * lane;
#pragma unroll
                for (int u = 0; u < 8; ++u) dst[u] = a[u][jj];
                wave_sync();
                T piv = col[j];
                T raw[8], cc[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) raw[u] = col[8 * bi + u];
#pragma unroll
                for (int v = 0; v < 8; ++v) cc[v] = col[8 * bk + v];
                const bool bad = !(piv > T(0));
                nbad += (bad && j < nb) ? 1 : 0;
                piv = bad ? Huge<T>::v : piv;
                const T rp = T(1) / piv;
                T m[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) m[u] = raw[u] * rp;
                // the next column first: it is published at the next step
                if (jj < 7) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u][jj + 1] -= m[u] * cc[jj + 1];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int v = 0; v < 8; ++v)
                        if (v != jj + 1) a[u][v] -= m[u] * cc[v];
                rps[j] = rp;
                pivs[j] = piv;
                __hip_atomic_store(&flag, j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (lane == 0 && nbad) atomicAdd(err, nbad);
        // L11[r][k] = l_rk sqrt(d_k) = raw_r(step k) / d_k * sqrt(d_k)
        const int r = lane;
        sdv[r] = sqrt(pivs[r]);
        wave_sync();
        T* trash = dump[0] + lane;
#pragma unroll
        for (int k = 0; k < kTile; ++k) {
            const T sk = sdv[k];
            const T f = (r > k) ? rps[k] * sk : T(0);
            T v = cbt[k * kTile + r] * f;
            v = (r == k) ? sk : v;
            T* dst = (r < nb && k < nb) ? P + (int64_t)k * c + r : trash;
            *dst = v;
        }
    } else {
        T x[8][8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < 8; ++v) x[u][v] = (8 * bi + u == 8 * bk + v) ? T(1) : T(0);
        for (int jb = 0; jb < 8; ++jb) {
            const bool lo = bi > jb, hi = bi >= jb;    // row 8bi+u > j for u <= jj / u > jj
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const int j = 8 * jb + jj;
                T* dx = (bi == jb) ? xrow + 8 * bk : dump[1] + 8 * lane;
#pragma unroll
                for (int v = 0; v < 8; ++v) dx[v] = x[jj][v];
                while (__hip_atomic_load(&flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) <= j)
                    __builtin_amdgcn_s_sleep(1);
                const T rp = rps[j];
                T m[8], xr[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const T mv = cbt[j * kTile + 8 * bi + u] * rp;
                    m[u] = ((u > jj) ? hi : lo) ? mv : T(0);
                }
#pragma unroll
                for (int v = 0; v < 8; ++v) xr[v] = xrow[8 * bk + v];
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int v = 0; v < 8; ++v) x[u][v] -= m[u] * xr[v];
                wave_sync();
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < 8; ++v) S[(8 * bk + v) * kS + 8 * bi + u] = x[u][v];
        wave_sync();
        const int r = lane;
        const T isd = T(1) / sqrt(pivs[r]);
        T* trash = dump[1] + lane;
#pragma unroll
        for (int k = 0; k < kTile; ++k) {
            const T v = S[k * kS + r] * ((r >= k) ? isd : T(0));
            T* dst = (r < nb && k < nb) ? Db + k * nb + r : trash;
            *dst = v;
        }
    }
}

// ------------------------------------------------------------- MFMA GEMM
// 4 waves, each a 32x32 quarter of a 64x64 tile built from 2x2 MFMA 16x16x4
// blocks.  Lane layout (measured on gfx950 by help_selftest_mfma): A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k; the result D[i][j] sits in lane
// j + 16(i%4), element i/4 for V_MFMA_F64_16X16X4_F64 and in lane
// j + 16(i/4), element i%4 for V_MFMA_F32_16X16X4_F32.
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;   // LDS and global address spaces
typedef __attribute__((address_space(1))) void glb_void;   // (GLOBAL_LOAD_LDS operands)

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double4_t acc_t;
    static __device__ inline acc_t fma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef float4_t acc_t;
    static __device__ inline acc_t fma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// acc += As[k][rows] * Bs[k][cols]^T over k in [0, kc), As/Bs k-major (ld 64)
template <typename T>
__device__ inline void mfma_tile(const T* As, const T* Bs, int kc, int lane, int wr, int wc,
                                 typename Mfma<T>::acc_t (&acc)[2][2]) {
    for (int kk = 0; kk < kc; kk += 4) {
        const int kr = kk + (lane >> 4);
        T a0 = As[kr * kTile + wr + (lane & 15)];
        T a1 = As[kr * kTile + wr + 16 + (lane & 15)];
        T b0 = Bs[kr * kTile + wc + (lane & 15)];
        T b1 = Bs[kr * kTile + wc + 16 + (lane & 15)];
        acc[0][0] = Mfma<T>::fma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::fma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::fma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::fma(a1, b1, acc[1][1]);
    }
}

// ------------------------- diagonal block, blocked form with MFMA updates
// Same result as k_potrf_diag (L11 = L_unit diag(sqrt(d)) in place, Dinv =
// diag(1/sqrt(d)) L_unit^-1), LDL^T by 4-column panels.  The 64x64 block and
// X = L_unit^-1 live in MFMA accumulator layout: wave w holds the 16x16
// tiles (w, 0..3) of each.  Block step kb, panel owner wave ow = kb/4:
//  * wave ow reads column block kb from S with the rows on the lanes and
//    factors it: pivot and in-panel values come from other lanes by
//    readlane, the multipliers m = a/d go to Mn (negated) and the unscaled
//    columns to Uf.  It then finishes the panel's 4 rows of X: they leave
//    its accumulators through XR, take the unit-lower 4x4 substitution with
//    the columns on the lanes, and go to XF (final, column-major) and back
//    to XR;
//  * barrier; every tile row below the panel applies A -= m U^T and
//    X -= m XR as one V_MFMA_F64_16X16X4_F64 per live tile (K = the 4 panel
//    columns) and stores its part of column block kb+1 to S; barrier.
// Two barriers per 4 pivots instead of per-pivot publishing, and the
// trailing updates run on the matrix cores.
// kTrsm: the task also carries row block rb - 1 below the diagonal block
// (every task of a panel factors the block redundantly; the task with
// rb <= 1 writes L11, Dinv and the pivot count).  Its 64 rows are loaded
// before the factorization and solved at the end as k_trsm_gemm does,
// X * Dinv^T on MFMA with Dinv taken from LDS: the same products and
// order, so the same values, without a launch of its own.
__device__ inline double read_lane(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

template <bool kTrsm>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_potrf_mfma(const PanelTask* __restrict__ tasks, DevSym sym, double* __restrict__ L,
                  double* __restrict__ Dinv, double* __restrict__ Lstage, int32_t* __restrict__ err) {
    typedef double T;
    typedef Mfma<double>::acc_t acc_t;
    constexpr int kS = kTile + 1;
    __shared__ T S[kTile * kS];
    __shared__ T Uf[kTile * kTile];        // column j before scaling (step j), Uf[j*64 + r]
    __shared__ T XF[kTile * kTile];        // final X, column-major
    __shared__ T Mn[kTile * 4];            // -m of the current panel, Mn[r*4 + jj]
    __shared__ T XR[4 * kTile];            // the current panel's rows of X
    __shared__ T rps[kTile], pivs[kTile], sdv[kTile];
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int nb = min(kTile, n - b);
    T* P = L + sym.loff[s] + (int64_t)b * c + b;
    T* Db = Dinv + dinv_block_offset(sym, s, b);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int lc = lane & 15, lk = lane >> 4;
    const bool owner = !kTrsm || tk.rb <= 1;
    const bool rows = kTrsm && tk.rb >= 1;
    const int32_t r0 = b + nb + (tk.rb - 1) * kTile;
    T* Lp = L + sym.loff[s];
    // the row block's entries (as k_trsm_gemm loads them), in flight
    // during the factorization
    T va[kTile * kTile / 256];
    if (rows) {
        const int32_t ri = min(r0 + lane, c - 1);
#pragma unroll
        for (int q = 0; q < kTile * kTile / 256; ++q) {
            const int kc = min(wave + 4 * q, nb - 1);
            va[q] = Lp[(int64_t)(b + kc) * c + ri];
        }
    }
    {
        // lower triangle, identity outside nb, zero above; clamped loads
        const int r = lane, rr = min(r, nb - 1);
        T v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = wave + 4 * q;
            v[q] = P[(int64_t)min(k, nb - 1) * c + rr];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = wave + 4 * q;
            S[k * kS + r] = (r < nb && k < nb) ? ((r >= k) ? v[q] : T(0)) : ((r == k) ? T(1) : T(0));
        }
    }
    __syncthreads();
    acc_t acc[4], xacc[4];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * wave + Mfma<double>::row(lane, r), col = 16 * tj + lc;
            acc[tj][r] = S[col * kS + row];
            xacc[tj][r] = (row == col) ? T(1) : T(0);
        }
    int nbad = 0;
    for (int kb = 0; kb < 16; ++kb) {
        if (wave == (kb >> 2)) {
            T p[4];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) p[jj] = S[(4 * kb + jj) * kS + lane];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = 4 * kb + jj;
                T d = read_lane(p[jj], j);
                T akj[4];
#pragma unroll
                for (int kk = jj + 1; kk < 4; ++kk) akj[kk] = read_lane(p[jj], 4 * kb + kk);
                const bool bad = !(d > T(0));
                nbad += (bad && j < nb) ? 1 : 0;
                d = bad ? Huge<T>::v : d;
                const T rp = T(1) / d;
                const T m = p[jj] * rp;
                Uf[j * kTile + lane] = p[jj];
                Mn[lane * 4 + jj] = -m;
#pragma unroll
                for (int kk = jj + 1; kk < 4; ++kk) p[kk] -= m * akj[kk];
                if (lane == 0) { rps[j] = rp; pivs[j] = d; }
            }
            // the panel's rows of X (rows 4(kb%4).. of this wave's tiles)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = Mfma<double>::row(lane, r);
                    if ((row >> 2) == (kb & 3)) XR[(row & 3) * kTile + 16 * tj + lc] = xacc[tj][r];
                }
            wave_sync();
            T xs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xs[u] = XR[u * kTile + lane];
#pragma unroll
            for (int u = 1; u < 4; ++u)
#pragma unroll
                for (int up = 0; up < u; ++up) xs[u] += Mn[(4 * kb + u) * 4 + up] * xs[up];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                XR[u * kTile + lane] = xs[u];
                XF[lane * kTile + 4 * kb + u] = xs[u];
            }
        }
        __syncthreads();
        if (16 * wave + 15 >= 4 * kb + 4) {
            const T aop = Mn[(16 * wave + lc) * 4 + lk];
            const T xop = (16 * wave + lc > 4 * kb + 3) ? aop : T(0);
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                if (16 * tj + 15 >= 4 * kb + 4) {          // live columns (uniform)
                    const T bop = Uf[(4 * kb + lk) * kTile + 16 * tj + lc];
                    acc[tj] = Mfma<double>::fma(aop, bop, acc[tj]);
                }
                if (tj <= wave && 16 * tj <= 4 * kb + 3) {  // nonzero part of X (uniform)
                    const T bx = XR[lk * kTile + 16 * tj + lc];
                    xacc[tj] = Mfma<double>::fma(xop, bx, xacc[tj]);
                }
            }
            if (kb + 1 < 16) {
                const int tj1 = (kb + 1) >> 2, jc = 4 * ((kb + 1) & 3);
                if (lc >= jc && lc < jc + 4) {
#pragma unroll
                    for (int tj = 0; tj < 4; ++tj)
                        if (tj == tj1)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                S[(16 * tj + lc) * kS + 16 * wave + Mfma<double>::row(lane, r)] = acc[tj][r];
                }
            }
        }
        __syncthreads();
    }
    if (owner && t == 0 && nbad) atomicAdd(err, nbad);
    if (t < kTile) sdv[t] = sqrt(pivs[t]);
    __syncthreads();
    // L11[r][k] = raw_r(step k) / d_k * sqrt(d_k); Dinv[r][k] = X[r][k] / sqrt(d_r)
    const int r = lane;
    const T isd = T(1) / sdv[r];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = wave + 4 * q;
        const T sk = sdv[k];
        const T f = (r > k) ? rps[k] * sk : T(0);
        T v = Uf[k * kTile + r] * f;
        v = (r == k) ? sk : v;
        if (kTrsm) {
            if (owner && r < nb && k < nb) Lstage[dinv_block_offset(sym, s, b) + k * nb + r] = v;
        } else {
            if (r < nb && k < nb) P[(int64_t)k * c + r] = v;
        }
        const T xv = XF[k * kTile + r] * ((r >= k) ? isd : T(0));
        if (owner && r < nb && k < nb) Db[k * nb + r] = xv;
    }
    if (!rows) return;   // uniform over the workgroup
    // row block: As[k][i] = X[i][k] in S's storage, Bs[k][j] = Dinv[j][k]
    // (the values just written to Db) in Uf's storage
    __syncthreads();
    T* As = S;
    T* Bs = Uf;
#pragma unroll
    for (int q = 0; q < kTile * kTile / 256; ++q) {
        const int k = wave + 4 * q;
        As[k * kTile + lane] = (k < nb && r0 + lane < c) ? va[q] : T(0);
        const T xv = XF[k * kTile + r] * ((r >= k) ? isd : T(0));
        Bs[k * kTile + lane] = (k < nb && lane < nb) ? xv : T(0);
    }
    __syncthreads();
    acc_t tacc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q) tacc[a][q] = acc_t{0, 0, 0, 0};
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    mfma_tile<T>(As, Bs, kTile, lane, wr, wc, tacc);
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q)
            for (int rr = 0; rr < 4; ++rr) {
                const int32_t i = wr + a * 16 + Mfma<T>::row(lane, rr);
                const int32_t j = wc + q * 16 + (lane & 15);
                if (r0 + i >= c || j >= nb) continue;
                Lp[(int64_t)(b + j) * c + r0 + i] = tacc[a][q][rr];
            }
}

// L11 block (s, b), nb x nb column-major at the Dinv offset of the block,
// back into the front's panel
__global__ __launch_bounds__(256) void k_copy_diag(const PanelTask* __restrict__ tasks, DevSym sym,
                                                   const double* __restrict__ Lstage, double* __restrict__ L) {
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int nb = min(kTile, n - b);
    const double* src = Lstage + dinv_block_offset(sym, s, b);
    double* P = L + sym.loff[s] + (int64_t)b * c + b;
    for (int idx = threadIdx.x; idx < nb * nb; idx += 256) {
        const int k = idx / nb, r = idx - k * nb;
        P[(int64_t)k * c + r] = src[idx];
    }
}

constexpr int kKc = 16;   // inner-dimension chunk staged in LDS

// C[row0:+64, col0:+64] -= P[rows, k0:k1] * P[cols, k0:k1]^T, lower part only,
// with P the L panel of the front.  The k chunks are double-buffered in LDS
// and the next chunk's global loads are in flight while the MFMAs consume
// the current one (one barrier per chunk).  At least 4 waves per SIMD are
// asked of the register allocation (the interior tiles' loads into LDS need
// no staging registers).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gemm(const GemmTask* __restrict__ tasks, DevSym sym,
                                              T* __restrict__ L, T* __restrict__ U, T* __restrict__ Uacc) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int kPer = kKc / 4;
    // the double-buffered k chunks, and after the product the 64x64
    // accumulator tile (column-major, leading dimension kTile + 1)
    constexpr int kCLd = kTile + 1;
    constexpr int kSmem = 4 * kKc * kTile > kTile * kCLd ? 4 * kKc * kTile : kTile * kCLd;
    __shared__ __attribute__((aligned(16))) T smem[kSmem];
    T (*As)[kKc * kTile] = reinterpret_cast<T (*)[kKc * kTile]>(smem);
    T (*Bs)[kKc * kTile] = reinterpret_cast<T (*)[kKc * kTile]>(smem + 2 * kKc * kTile);
    const GemmTask tk = tasks[blockIdx.x];
    const int32_t n = tk.n;
    const int32_t c = tk.c;
    const int32_t cu = c - n;
    const T* Lp = L + tk.loff;
    T* Lw = L + tk.loff;
    T* Up = U + tk.uoff;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    // update-matrix tile continued from an earlier panel: reload the parked
    // accumulator (entries outside the tile's lower part read element 0)
    T* Ua = Uacc ? Uacc + tk.uoff : nullptr;
    if (tk.mode == kGemmNext || tk.mode == kGemmLast) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int32_t i = tk.row0 + wr + a * 16 + Mfma<T>::row(lane, r);
                    const int32_t j = tk.col0 + wc + b * 16 + (lane & 15);
                    const bool ok = i < c && j < tk.col_end && i >= j;
                    acc[a][b][r] = Ua[ok ? (int64_t)(j - n) * cu + (i - n) : 0];
                }
    }
    if (tk.row0 + kTile <= c && tk.col0 + kTile <= c) {   // uniform over the workgroup
        // Interior tile: the chunks go from HBM straight into LDS
        // (GLOBAL_LOAD_LDS_DWORDX4: lane l of a wave lands 16 bytes at the
        // wave's LDS base + 16 l).  Wave w moves columns 4w .. 4w+3 of the
        // chunk, 64 / kV lanes per column with kV rows each (kV = 4 FP32, 2
        // FP64: one or two loads per operand), which is the k-major chunk
        // layout.  Columns past the panel read its last column and are
        // zeroed in LDS once they have landed (last chunk only).  The chunks
        // hold the same values as the register-staged forms.
        constexpr int kV = 16 / (int)sizeof(T);
        constexpr int kLc = kTile / kV;        // lanes per column
        constexpr int kCpi = 64 / kLc;         // columns per load
        const int rr = (lane % kLc) * kV;
        const T* pa = Lp + tk.row0 + rr;
        const T* pb = Lp + tk.col0 + rr;
        auto issue = [&](int32_t kb, int b) {
#pragma unroll
            for (int q = 0; q < 4 / kCpi; ++q) {
                const int k = 4 * wave + q * kCpi;
                const int64_t o = (int64_t)min(kb + k + lane / kLc, tk.k1 - 1) * c;
                __builtin_amdgcn_global_load_lds((const glb_void*)(pa + o), (lds_void*)&As[b][k * kTile], 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const glb_void*)(pb + o), (lds_void*)&Bs[b][k * kTile], 16, 0, 0);
            }
        };
        issue(tk.k0, 0);
        int buf = 0;
        for (int32_t kb = tk.k0; kb < tk.k1; kb += kKc, buf ^= 1) {
            __syncthreads();   // this chunk has landed; the other buffer's readers are done
            if (kb + kKc < tk.k1) {
                issue(kb + kKc, buf ^ 1);
            } else if (tk.k1 - kb < kKc) {
                for (int idx = tid; idx < kKc * kTile; idx += 256)
                    if (kb + idx / kTile >= tk.k1) { As[buf][idx] = T(0); Bs[buf][idx] = T(0); }
                __syncthreads();
            }
            mfma_tile<T>(As[buf], Bs[buf], kKc, lane, wr, wc, acc);
        }
    } else if constexpr (sizeof(T) == 4) {
      {
        // FP32: each thread moves 4 consecutive rows of one column of the
        // chunk with one 16-byte load (4-byte aligned: the column stride c is
        // arbitrary, and dword alignment is what multi-dword global loads
        // need); rows past the front read zeros through clamped loads.  The
        // LDS chunks hold the same values as the scalar form below.
        typedef float vload_t __attribute__((ext_vector_type(4), aligned(4)));
        const int vr = (tid & 15) * 4, vk = tid >> 4;   // 16 columns, 64 rows
        const int32_t ia = tk.row0 + vr, ib = tk.col0 + vr;
        const bool afull = ia + 4 <= c, bfull = ib + 4 <= c;
        float4_t ra, rb;
        auto rows4 = [&](const T* col, int32_t i0, bool full) {
            if (full) {
                const vload_t v = *reinterpret_cast<const vload_t*>(col + i0);
                return float4_t{v[0], v[1], v[2], v[3]};
            }
            float4_t v;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const T x = col[min(i0 + e, c - 1)];
                v[e] = (i0 + e < c) ? x : T(0);
            }
            return v;
        };
        auto load = [&](int32_t kb) {
            const T* col = Lp + (int64_t)min(kb + vk, tk.k1 - 1) * c;
            ra = rows4(col, ia, afull);
            rb = rows4(col, ib, bfull);
        };
        load(tk.k0);
        int buf = 0;
        for (int32_t kb = tk.k0; kb < tk.k1; kb += kKc, buf ^= 1) {
            const bool kin = kb + vk < tk.k1;
            *reinterpret_cast<float4_t*>(&As[buf][vk * kTile + vr]) = kin ? ra : float4_t{0, 0, 0, 0};
            *reinterpret_cast<float4_t*>(&Bs[buf][vk * kTile + vr]) = kin ? rb : float4_t{0, 0, 0, 0};
            __syncthreads();
            if (kb + kKc < tk.k1) load(kb + kKc);
            mfma_tile<T>(As[buf], Bs[buf], kKc, lane, wr, wc, acc);
        }
      }
    } else {
        const int li = tid & 63, lk = tid >> 6;
        const int32_t ia = tk.row0 + li, ib = tk.col0 + li;
        const bool ain = ia < c, bin = ib < c;
        const T* pa = Lp + min(ia, c - 1);
        const T* pb = Lp + min(ib, c - 1);
        T ra[kPer], rb[kPer];
        auto load = [&](int32_t kb) {
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const int32_t kg = min(kb + lk + 4 * r, tk.k1 - 1);
                ra[r] = pa[(int64_t)kg * c];
                rb[r] = pb[(int64_t)kg * c];
            }
        };
        load(tk.k0);
        int buf = 0;
        for (int32_t kb = tk.k0; kb < tk.k1; kb += kKc, buf ^= 1) {
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const int k = lk + 4 * r;
                const bool kin = kb + k < tk.k1;
                As[buf][k * kTile + li] = (kin && ain) ? ra[r] : T(0);
                Bs[buf][k * kTile + li] = (kin && bin) ? rb[r] : T(0);
            }
            __syncthreads();
            if (kb + kKc < tk.k1) load(kb + kKc);
            mfma_tile<T>(As[buf], Bs[buf], kKc, lane, wr, wc, acc);
        }
    }
    if (tk.mode == kGemmFirst || tk.mode == kGemmNext) {
        // park the accumulator for the next panel
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int32_t i = tk.row0 + wr + a * 16 + Mfma<T>::row(lane, r);
                    const int32_t j = tk.col0 + wc + b * 16 + (lane & 15);
                    if (i < c && j < tk.col_end && i >= j) Ua[(int64_t)(j - n) * cu + (i - n)] = acc[a][b][r];
                }
        return;
    }
    // read-modify-write of C.  The accumulator tile goes through LDS so
    // that the lanes run along the rows of one column per access (64
    // consecutive entries per load or store); each thread takes one row of
    // 16 columns and issues its 16 loads before its stores.  A task's tile
    // lies in L's columns (col_end <= n) or in the update matrix
    // (col0 >= n), so one tile origin and 32-bit offsets address it; entries
    // outside the lower part or the front read the origin and are not
    // written back.
    __syncthreads();   // every wave is done with the k chunks
    T* Cs = smem;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int32_t ti = wr + a * 16 + Mfma<T>::row(lane, r);
                const int32_t tj = wc + b * 16 + (lane & 15);
                Cs[tj * kCLd + ti] = acc[a][b][r];
            }
    __syncthreads();
    const bool inU = tk.col0 >= n;
    const int32_t ld = inU ? cu : c;
    T* Ct = inU ? Up + (int64_t)(tk.col0 - n) * cu + (tk.row0 - n) : Lw + (int64_t)tk.col0 * c + tk.row0;
    const int32_t ti = tid & 63, tj0 = (tid >> 6) * 16;
    const int32_t i = tk.row0 + ti;
    T cv[16];
    bool okv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int32_t tj = tj0 + q, j = tk.col0 + tj;
        okv[q] = i < c && j < tk.col_end && i >= j;
        cv[q] = Ct[okv[q] ? tj * ld + ti : 0];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q)
        if (okv[q]) Ct[(tj0 + q) * ld + ti] = cv[q] - Cs[(tj0 + q) * kCLd + ti];
}

// X := X L_bb^-T for the 64 rows of block rb-1 below the diagonal block,
// computed as X * (L_bb^-1)^T on MFMA in place.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_gemm(const PanelTask* __restrict__ tasks, DevSym sym,
                                                   T* __restrict__ L, const T* __restrict__ Dinv) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[kTile * kTile];
    __shared__ T Bs[kTile * kTile];
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t nb = min(kTile, n - b);
    const int32_t r0 = b + nb + (tk.rb - 1) * kTile;
    T* Lp = L + sym.loff[s];
    const T* Db = Dinv + dinv_block_offset(sym, s, b);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    // As[k][i] = X[i][k] = P[r0+i][b+k];  Bs[k][j] = Dinv[j][k] (Dinv column-major:
    // Db[k*nb + j] = Dinv[j][k]).  Every load is issued before the first LDS
    // store: out-of-range elements read a clamped in-range address and are
    // zeroed afterwards, so no branch separates the loads.
    constexpr int kPer = kTile * kTile / 256;
    T va[kPer], vb[kPer];
    {
        const int i = tid & 63;
        const int32_t ri = min(r0 + i, c - 1);
        const int ib = min(i, nb - 1);
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int k = (tid >> 6) + 4 * t;
            const int kc = min(k, nb - 1);
            va[t] = Lp[(int64_t)(b + kc) * c + ri];
            vb[t] = Db[kc * nb + ib];
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int k = (tid >> 6) + 4 * t;
            As[k * kTile + i] = (k < nb && r0 + i < c) ? va[t] : T(0);
            Bs[k * kTile + i] = (k < nb && i < nb) ? vb[t] : T(0);
        }
    }
    __syncthreads();
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q) acc[a][q] = acc_t{0, 0, 0, 0};
    mfma_tile<T>(As, Bs, kTile, lane, wr, wc, acc);
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q)
            for (int r = 0; r < 4; ++r) {
                int32_t i = wr + a * 16 + Mfma<T>::row(lane, r);
                int32_t j = wc + q * 16 + (lane & 15);
                if (r0 + i >= c || j >= nb) continue;
                Lp[(int64_t)(b + j) * c + r0 + i] = acc[a][q][r];
            }
}

// acc += As[k][rows] * Bs[k][cols]^T as mfma_tile, with a padded Bs (leading
// dimension ldb) so that column-major blocks can be staged with coalesced loads
template <typename T>
__device__ inline void mfma_tile_ld(const T* As, const T* Bs, int ldb, int lane, int wr, int wc,
                                    typename Mfma<T>::acc_t (&acc)[2][2]) {
    for (int kk = 0; kk < kTile; kk += 4) {
        const int kr = kk + (lane >> 4);
        T a0 = As[kr * kTile + wr + (lane & 15)];
        T a1 = As[kr * kTile + wr + 16 + (lane & 15)];
        T b0 = Bs[kr * ldb + wc + (lane & 15)];
        T b1 = Bs[kr * ldb + wc + 16 + (lane & 15)];
        acc[0][0] = Mfma<T>::fma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::fma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::fma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::fma(a1, b1, acc[1][1]);
    }
}

// One step of the right-looking inverse of a wide supernode's L11, task
// (i, j, k = j + t):
//   t = 0:  Xk = Dinv_j                       and S_ij  = L_ij Xk   (i > j)
//   t > 0:  Xk = -Dinv_k S_kj  (= X_kj);  i == k: store X_kj,
//                                         i >  k: S_ij += L_ik Xk
// Every task recomputes Xk from S_kj, which no task of the launch writes, so
// the tasks are independent; S_ij gains its terms in k order.  Blocks are
// column-major 64x64; rows and columns past the supernode are zero.
template <typename T>
__global__ __launch_bounds__(256) void k_xinv(const XTask* __restrict__ tasks, DevSym sym,
                                              const T* __restrict__ L, const T* __restrict__ Dinv,
                                              T* __restrict__ X, T* __restrict__ Sb) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int kPer = kTile * kTile / 256;
    constexpr int kLdb = kTile + 1;
    __shared__ T As[kTile * kTile];
    __shared__ T Bs[kTile * kLdb];
    const XTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, i = tk.i, j = tk.j, k = tk.k;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t nbi = min(kTile, n - kTile * i), nbk = min(kTile, n - kTile * k);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = tid & 63, lk = tid >> 6;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q) acc[a][q] = acc_t{0, 0, 0, 0};
    if (k == j) {
        // Bs[kk][q] = Dinv_j[kk][q] (a full block: j < i)
        const T* Dj = Dinv + sym.dinvoff[s] + (int64_t)j * 4096;
        T vb[kPer];
#pragma unroll
        for (int t = 0; t < kPer; ++t) vb[t] = Dj[(lk + 4 * t) * kTile + li];   // element (li, q)
#pragma unroll
        for (int t = 0; t < kPer; ++t) Bs[li * kLdb + lk + 4 * t] = vb[t];
    } else {
        // Xk = -Dinv_k S_kj: As[kk][r] = Dinv_k[r][kk] (ld nbk), Bs[kk][q] = S_kj[kk][q]
        const T* Dk = Dinv + sym.dinvoff[s] + (int64_t)k * 4096;
        const T* Skj = Sb + xblock_offset(sym.xoff, s, k, j);
        const int rr = min(li, nbk - 1);
        T va[kPer], vb[kPer];
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int kk = lk + 4 * t;
            va[t] = Dk[min(kk, nbk - 1) * nbk + rr];
            vb[t] = Skj[kk * kTile + li];                                    // element (li, kk)
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int kk = lk + 4 * t;
            As[kk * kTile + li] = (li < nbk && kk < nbk) ? va[t] : T(0);
            Bs[li * kLdb + kk] = vb[t];
        }
        __syncthreads();
        mfma_tile_ld<T>(As, Bs, kLdb, lane, wr, wc, acc);
        __syncthreads();
        T* Xkj = X + xblock_offset(sym.xoff, s, k, j);
        for (int a = 0; a < 2; ++a)
            for (int q = 0; q < 2; ++q)
                for (int r = 0; r < 4; ++r) {
                    const int rw = wr + a * 16 + Mfma<T>::row(lane, r);
                    const int cl = wc + q * 16 + (lane & 15);
                    const T v = -acc[a][q][r];
                    Bs[rw * kLdb + cl] = v;       // Xk[rw][cl] feeds the next product
                    if (i == k) Xkj[cl * kTile + rw] = v;
                    acc[a][q][r] = T(0);
                }
        if (i == k) return;
    }
    // S_ij (+)= L_ik Xk: As[kk][r] = L_ik[r][kk]
    {
        const T* Lp = L + sym.loff[s];
        const int32_t row = kTile * i + min(li, nbi - 1);
        T va[kPer];
#pragma unroll
        for (int t = 0; t < kPer; ++t) va[t] = Lp[(int64_t)(kTile * k + lk + 4 * t) * c + row];
#pragma unroll
        for (int t = 0; t < kPer; ++t) As[(lk + 4 * t) * kTile + li] = (li < nbi) ? va[t] : T(0);
    }
    __syncthreads();
    mfma_tile_ld<T>(As, Bs, kLdb, lane, wr, wc, acc);
    T* Sij = Sb + xblock_offset(sym.xoff, s, i, j);
    // all loads of the previous partial sums before the stores (at t = 0
    // the block is written fresh; its stale contents are read and dropped)
    T prev[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rw = wr + a * 16 + Mfma<T>::row(lane, r);
                const int cl = wc + q * 16 + (lane & 15);
                prev[a][q][r] = Sij[cl * kTile + rw];
            }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rw = wr + a * 16 + Mfma<T>::row(lane, r);
                const int cl = wc + q * 16 + (lane & 15);
                const T pv = (k == j) ? T(0) : prev[a][q][r];
                Sij[cl * kTile + rw] = pv + acc[a][q][r];
            }
}

// one wave: lane l feeds a[l], b[l]; writes its 4 result elements
template <typename T>
__global__ void k_mfma_probe(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int l = threadIdx.x;
    acc_t acc = acc_t{0, 0, 0, 0};
    acc = Mfma<T>::fma(a[l], b[l], acc);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
}

}  // namespace

template <typename T>
void launch_mfma_probe(const T* a, const T* b, T* out, hipStream_t st) {
    help::launch_kernel(k_mfma_probe<T>, 1, 64, 0, st, a, b, out);
}
template void launch_mfma_probe<double>(const double*, const double*, double*, hipStream_t);
template void launch_mfma_probe<float>(const float*, const float*, float*, hipStream_t);

template <typename T>
void launch_small_fronts(const int32_t* list, int32_t count, DevSym sym, T* L, T* U, T* Dinv,
                         int32_t* err, int32_t lds_elems, hipStream_t st) {
    if (count > 0)
        help::launch_kernel(k_small_fronts<T>, count, 256, (size_t)lds_elems * sizeof(T), st, list, sym, L, U, Dinv, err);
}
template <typename T>
void launch_small_wave(const int32_t* list, int32_t count, DevSym sym, T* L, T* U, T* Dinv, int32_t* err,
                       int32_t lds_elems, int32_t lanes, hipStream_t st) {
    if (count <= 0) return;
    if (lanes == 32)
        help::launch_kernel(k_small_wave<T, 32>, (count + 7) / 8, 256, (size_t)8 * lds_elems * sizeof(T), st,
                            list, count, lds_elems, sym, L, U, Dinv, err);
    else
        help::launch_kernel(k_small_wave<T, 64>, (count + 3) / 4, 256, (size_t)4 * lds_elems * sizeof(T), st,
                            list, count, lds_elems, sym, L, U, Dinv, err);
}
template <typename T>
void launch_extend_add(const EaTask* tasks, int32_t count, const EaRef* refs, DevSym sym, T* L, T* U,
                       int32_t depth, hipStream_t st) {
    if (count <= 0) return;
    if (depth == 8) help::launch_kernel(k_extend_add<T, 8>, count, 256, 0, st, tasks, refs, sym, L, U);
    else help::launch_kernel(k_extend_add<T, 4>, count, 256, 0, st, tasks, refs, sym, L, U);
}
template <typename T>
void launch_gemm(const GemmTask* tasks, int32_t count, DevSym sym, T* L, T* U, T* Uacc, hipStream_t st) {
    if (count > 0) help::launch_kernel(k_gemm<T>, count, 256, 0, st, tasks, sym, L, U, Uacc);
}

template <typename T>
void launch_potrf_diag(const PanelTask* tasks, int32_t count, DevSym sym, T* L, T* Dinv, int32_t* err,
                       hipStream_t st) {
    if (count <= 0) return;
    // FP64: the blocked MFMA form; FP32 keeps the two-wave register-blocked form
    if constexpr (std::is_same<T, double>::value)
        help::launch_kernel(k_potrf_mfma<false>, count, 256, 0, st, tasks, sym, L, Dinv, (double*)nullptr, err);
    else
        help::launch_kernel(k_potrf_diag<T>, count, 128, 0, st, tasks, sym, L, Dinv, err);
}
void launch_panel_fp64(const PanelTask* tasks, int32_t count, DevSym sym, double* L, double* Dinv,
                       double* Lstage, int32_t* err, hipStream_t st) {
    if (count > 0)
        help::launch_kernel(k_potrf_mfma<true>, count, 256, 0, st, tasks, sym, L, Dinv, Lstage, err);
}
void launch_copy_diag_fp64(const PanelTask* tasks, int32_t count, DevSym sym, const double* Lstage,
                           double* L, hipStream_t st) {
    if (count > 0) help::launch_kernel(k_copy_diag, count, 256, 0, st, tasks, sym, Lstage, L);
}
template <typename T>
void launch_trsm_rows(const PanelTask* tasks, int32_t count, DevSym sym, T* L, const T* Dinv,
                      hipStream_t st) {
    if (count > 0) help::launch_kernel(k_trsm_gemm<T>, count, 256, 0, st, tasks, sym, L, Dinv);
}
template <typename T>
void launch_xinv(const XTask* tasks, int32_t count, DevSym sym, const T* L, const T* Dinv, T* X, T* S,
                 hipStream_t st) {
    if (count > 0) help::launch_kernel(k_xinv<T>, count, 256, 0, st, tasks, sym, L, Dinv, X, S);
}

#define HELP_INST(T)                                                                                \
    template void launch_small_fronts<T>(const int32_t*, int32_t, DevSym, T*, T*, T*, int32_t*,      \
                                         int32_t, hipStream_t);                                     \
    template void launch_small_wave<T>(const int32_t*, int32_t, DevSym, T*, T*, T*, int32_t*,        \
                                       int32_t, int32_t, hipStream_t);                              \
    template void launch_extend_add<T>(const EaTask*, int32_t, const EaRef*, DevSym, T*, T*, int32_t, hipStream_t); \
    template void launch_gemm<T>(const GemmTask*, int32_t, DevSym, T*, T*, T*, hipStream_t);           \
    template void launch_potrf_diag<T>(const PanelTask*, int32_t, DevSym, T*, T*, int32_t*,        \
                                       hipStream_t);                                                \
    template void launch_trsm_rows<T>(const PanelTask*, int32_t, DevSym, T*, const T*, hipStream_t);     \
    template void launch_xinv<T>(const XTask*, int32_t, DevSym, const T*, const T*, T*, T*, hipStream_t);
HELP_INST(double)
HELP_INST(float)
#undef HELP_INST

}  // namespace help
