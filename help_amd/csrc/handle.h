// Private definitions shared by the backend translation units (the handle,
// device buffers, error plumbing).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"
#include "help_mi355x.h"
#include "mpnn.h"
#include "pcg.h"
#include "symbolic.h"

namespace help {
// last help_gpu_create error (no handle to hold it)
extern thread_local std::string g_create_error;

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& s, int c) : std::runtime_error(s), code(c) {}
};

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_), HELP_ERR_HIP);  \
    } while (0)

template <typename T>
struct DBuf {   // owning device buffer
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t cnt) {
        free();
        n = cnt;
        if (cnt) HIPCHK(hipMalloc((void**)&p, cnt * sizeof(T)));
    }
    void upload(const std::vector<T>& v, hipStream_t st) {
        alloc(v.size());
        if (n) HIPCHK(hipMemcpyAsync(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { free(); }
};

// kernel classes (profiling index = HELP_K_* in the public header)
enum OpKind : int32_t {
    OP_SMALL = HELP_K_SMALL_FRONTS, OP_EA = HELP_K_EXTEND_ADD, OP_POTRF = HELP_K_POTRF,
    OP_TRSM = HELP_K_TRSM, OP_GEMM = HELP_K_GEMM,
    OP_FWD = HELP_K_FWD, OP_FWD_GEMV = HELP_K_FWD_GEMV, OP_BWD_GEMV = HELP_K_BWD_GEMV,
    OP_BWD = HELP_K_BWD, OP_XINV = HELP_K_XINV, OP_FWD_XINV = HELP_K_FWD_XINV,
    OP_BWD_XINV = HELP_K_BWD_XINV
};
struct Op {
    OpKind kind;
    int32_t off, count;
    double work;       // algorithmic flops (factor classes) or bytes (the others)
    int32_t lds = 0;   // dynamic LDS elements per workgroup (small fronts)
    int32_t ntiny = 0; // solves: leading supernodes of the list solved one per wave
    int32_t xoff = 0, xcount = 0;   // solves: the level's wide-supernode block tasks;
                                    // FP64 factor: the phase's fused panel tasks
                                    // (POTRF; its TRSM op then has fused set)
    double work2 = 0;  // flops of the fused panel launch
    int32_t wave = 0;  // lanes per front or supernode of the one-wave forms (64 or 32):
                       // small fronts (0 = workgroup kernel) and the tiny supernodes
                       // of a solve level
    bool fused = false;  // TRSM: the phase's panels ran in the fused POTRF launch
    int32_t depth = 4;   // extend-add: child rows per lane in flight (4, or 8 for large fronts)
};

inline const char* const kKindName[HELP_K_COUNT] = {
    "k_small_fronts", "k_extend_add", "k_potrf_diag", "k_trsm_rows", "k_gemm",
    "k_fwd", "k_fwd_gemv", "k_bwd_gemv", "k_bwd", "k_adat", "k_rhs", "k_backsub",
    "pcg_build", "pcg_solve", "k_xinv", "k_fwd_xinv", "k_bwd_xinv"};

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace help

using help::AdatPattern;
using help::CscMatrix;
using help::CsrMatrix;
using help::DBuf;
using help::DevSym;
using help::EaTask;
using help::GemmTask;
using help::HipError;
using help::Op;
using help::PanelTask;
using help::XTask;
using help::SolveTask;
using help::Supernodal;

struct help_gpu {
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    help_opts opts{};
    CscMatrix A;
    CsrMatrix R;
    AdatPattern P;
    Supernodal S;
    std::vector<int64_t> slot_dest, diagpos;
    std::vector<int32_t> colof, csrcol;
    std::vector<double> csrval;
    std::vector<int8_t> sbig;
    // schedules
    std::vector<Op> fops, sops_fwd, sops_bwd;
    std::vector<int32_t> list_small, list_fwd, list_bwd;
    std::vector<EaTask> ea;
    std::vector<help::EaRef> earef;
    std::vector<PanelTask> potrf, trsm, panel;
    std::vector<PanelTask> pcopy;   // diagonal blocks factored by fused panel launches
    std::vector<GemmTask> gemm;
    std::vector<SolveTask> fgemv, bgemv;
    // explicit inverse of L11 for supernodes wider than one 64-column block
    bool use_xinv = false;
    bool use_uacc = false;   // some update matrix is built over several panels (Uacc)
    std::vector<int64_t> xoff;             // nsuper+1 offsets of the inverse blocks
    // update-matrix placement: uplan[s] = element offset of front s's update
    // matrix (and of its parked accumulator) in the shared update buffer,
    // uplan[nsuper] = the buffer's length (fronts whose lifetimes do not
    // overlap share memory)
    std::vector<int64_t> uplan;
    std::vector<XTask> xtask;
    std::vector<SolveTask> fxs, bxs;       // row blocks (forward) / column blocks (backward)
    std::string err;
    bool device = false;
    bool factored = false;
    bool use_fp32 = false;     // FP32 mode requested
    bool factor_fp32 = false;  // precision of the current factor
    bool fp32_off = false;     // FP32 abandoned (refinement failure or its cost)
    int64_t cyc_solves = 0, cyc_steps = 0;   // FP32 solves / refinement steps since the last renew
    int32_t last_tiny = 0;
    // device state
    hipStream_t st = nullptr;
    hipEvent_t ev[6] = {};
    std::vector<int64_t> dinvoff;
    DBuf<int64_t> d_colptr, d_rowptr, d_pairptr, d_dest, d_diagpos, d_loff, d_uoff, d_rowoff, d_reloff,
        d_dinvoff;
    DBuf<int32_t> d_rowind, d_colind, d_pair_p, d_pair_q, d_colof, d_iperm, d_sfirst, d_snrow,
        d_rowidx, d_relmap, d_childptr, d_child, d_list_small, d_list_fwd, d_list_bwd, d_err;
    DBuf<double> d_val, d_csrval, d_d, d_dinv, d_b1, d_b2, d_x1, d_x2, d_rp, d_W, d_t, d_tn, d_r,
        d_c, d_z, d_p, d_q, d_part, d_scal, d_slotvals;
    DBuf<double> d_wnz;   // A_ij / d_j per nonzero (the assembly's first factor)
    DBuf<double> d_L64, d_Dinv64;
    // update matrices and parked accumulators: scratch of the factorization,
    // one buffer each for both precisions (FP32 uses its first half)
    DBuf<double> d_Ubuf, d_Uaccbuf;
    DBuf<double> d_Lstage64;   // L11 of the fused panels (Dinv layout) until the final copy
    // FP32 mode: symmetric diagonal scaling of the factored matrix
    std::vector<int32_t> slot_row;
    DBuf<int32_t> d_slot_row, d_slot_col;
    DBuf<double> d_scale;
    DBuf<float> d_L32, d_Dinv32;
    DBuf<int8_t> d_big;
    DBuf<EaTask> d_ea;
    DBuf<help::EaRef> d_earef;
    DBuf<PanelTask> d_potrf, d_trsm, d_panel, d_pcopy;
    DBuf<GemmTask> d_gemm;
    DBuf<SolveTask> d_fgemv, d_bgemv, d_fxs, d_bxs;
    DBuf<XTask> d_xtask;
    DBuf<int64_t> d_xoff;
    DBuf<float> d_X32, d_S32;              // inverse blocks and their partial sums
    int64_t w_len = 1;            // update-vector entries per right-hand side
    DBuf<double> d_X64, d_S64, d_y;
    DevSym sym{};
    // side stream of the factorization: the inverse blocks of a finished
    // level are formed there while the main stream factors the next levels
    // (fork: main -> side, join: side -> main)
    hipStream_t st_side = nullptr;
    hipEvent_t ev_side[2] = {nullptr, nullptr};
    // the factorization's launch sequence captured once per precision
    // (opts.factor_graph): [0] FP64, [1] FP32
    hipGraphExec_t fgraph[2] = {nullptr, nullptr};
    // pinned host words for results read after the solve's final
    // synchronization (the residual norms of an unrefined solve)
    double* h_pin = nullptr;
    bool rel_deferred = false;
    // profiling
    bool prof = false;
    struct KStat { double ms = 0, work = 0; int64_t launches = 0; } kstat[HELP_K_COUNT];
    KStat kstat32[HELP_K_COUNT];   // the part of kstat run on the FP32 factor
    bool prof_fp32 = false;        // the launches being recorded use the FP32 factor
    struct Pending { int32_t kind; double work; hipEvent_t a, b; bool fp32; };
    std::vector<hipEvent_t> evpool;
    std::vector<Pending> pending;
    double work_adat = 0, work_rhs = 0, work_backsub = 0;
    // row partition of A for products with heavy rows (all modes)
    DBuf<help::RowChunk> d_chunks;
    DBuf<int32_t> d_hrow, d_hptr, d_mrow;
    DBuf<double> d_rpart;
    help::RowSplitDev rows;
    int32_t max_row_nnz = 0;
    // matrix-free PCG mode (pcg.hip)
    help::PcgState* pcg = nullptr;
    bool use_pcg = false;
    bool learned = false;          // PCG started from the network's guess
    help_mpnn* net = nullptr;      // not owned

    ~help_gpu() {
        if (st) { (void)hipStreamSynchronize(st); }
        if (st_side) { (void)hipStreamSynchronize(st_side); }
        for (auto& g : fgraph)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto& e : ev_side)
            if (e) (void)hipEventDestroy(e);
        if (st_side) (void)hipStreamDestroy(st_side);
        if (h_pin) (void)hipHostFree(h_pin);
        help::pcg_free(pcg);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : evpool) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace help {
// PCG mode (pcg.hip)
void pcg_setup(help_gpu* h);
void pcg_renew(help_gpu* h);
int pcg_solve(help_gpu* h, const double* rhs, const double* x0, double* x2, double& rel);
// algorithmic HBM bytes: one partial-Cholesky build, one CG iteration
double pcg_build_bytes(const help_gpu* h);
double pcg_iteration_bytes(const help_gpu* h);
void pcg_download(help_gpu* h, int32_t* K, int32_t* perm, double* D, double* S, double* L);
void pcg_precond(help_gpu* h, const double* v, double* out);
void pcg_precond_noalloc(help_gpu* h, const double* v, double* out);
// warm-start network forward on device vectors (mpnn.hip)
void mpnn_predict_device(help_mpnn* net, help_gpu* h, const double* d, const double* r, double* x0);
}  // namespace help
