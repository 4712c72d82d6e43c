// Host C++ of the backend under AddressSanitizer and UndefinedBehaviour-
// Sanitizer (built and run by tests/test_host_sanitizers.py, CPU only):
// the instance generators and file formats (generators.cpp), the symbolic
// analysis with both orderings and the assembly maps (symbolic.cpp), and the
// HSD driver restatement (hsd.cpp) over the CPU oracle backend, on small
// instances of every generator, bounded and free columns included.  Exits
// non-zero on a wrong result; the sanitizers abort on a memory or UB error.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "help_mi355x.h"
#include "symbolic.h"

#include "oracle.h"

using namespace help;

static int fails = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) { std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); ++fails; } \
    } while (0)

static CscMatrix csc_of(const help_lp& lp) {
    CscMatrix A;
    A.m = lp.m;
    A.n = lp.n;
    A.colptr.assign(lp.colptr, lp.colptr + lp.n + 1);
    A.rowind.assign(lp.rowind, lp.rowind + lp.nnz);
    A.val.assign(lp.val, lp.val + lp.nnz);
    return A;
}

static void symbolic_checks(const help_lp& lp, int ordering) {
    CscMatrix A = csc_of(lp);
    CsrMatrix R = csc_to_csr(A);
    AdatPattern P = adat_pattern(A, R, true);
    std::vector<int64_t> ap;
    std::vector<int32_t> aj;
    adat_adjacency(A, R, ap, aj);
    // the adjacency is the off-diagonal upper pattern mirrored, rows sorted
    int64_t offdiag = 0;
    for (int32_t i = 0; i < A.m; ++i) {
        for (int64_t t = ap[i] + 1; t < ap[i + 1]; ++t) CHECK(aj[t - 1] < aj[t]);
        for (int64_t s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
            const int32_t k = P.colind[s];
            if (k == i) continue;
            ++offdiag;
            CHECK(std::binary_search(aj.begin() + ap[i], aj.begin() + ap[i + 1], k));
            CHECK(std::binary_search(aj.begin() + ap[k], aj.begin() + ap[k + 1], i));
        }
    }
    CHECK(ap[A.m] == 2 * offdiag);
    std::vector<int32_t> perm;
    if (ordering == 1) perm = nd_order(A.m, ap, aj, 64);
    RelaxParams rp;
    Supernodal S = analyze(A.m, ap, aj, perm, rp);
    std::vector<int32_t> seen(A.m, 0);
    for (int32_t v : S.perm) if (v >= 0 && v < A.m) seen[v]++;
    for (int32_t i = 0; i < A.m; ++i) CHECK(seen[i] == 1);
    std::vector<int32_t> walk = colcounts_walk(A.m, ap, aj, S.perm, S.iperm, S.parent);
    CHECK(walk == S.colcount);
    std::vector<int64_t> dest, diag;
    std::vector<int32_t> colof;
    std::vector<double> csrval;
    assembly_maps(S, P, A, R, dest, diag, colof, csrval);
    CHECK(dest.size() == P.colind.size());
    for (int64_t x : dest) CHECK(x >= 0 && x < S.loff[S.nsuper]);
}

static void hsd_check(const help_lp& lp, int32_t flags) {
    or_backend* ob = or_backend_new(1e-5, nullptr, 0);
    help_backend be{ob, reinterpret_cast<int (*)(void*, const help_lp*)>(or_be_construct), or_be_destruct,
                    or_be_renew, or_be_solve};
    help_hsd_result res{};
    int st = help_hsd_solve(&lp, &be, 0, flags, nullptr, nullptr, &res);
    or_backend_free(ob);
    CHECK(st == HELP_OK);
    CHECK(res.status == HELP_HSD_OK);
    CHECK(res.rel_gap <= 1e-8);
}

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    std::vector<help_gen_params> gens;
    help_gen_params g{};
    g.kind = HELP_GEN_RANDOM; g.seed = 81; g.m = 200; g.nb = 400; g.k = 4;
    gens.push_back(g);
    g.ub_cols = 80; g.free_cols = 3; g.seed = 1013; g.m = 300; g.nb = 600;   // bounded + free, reaches OPTIMAL
    gens.push_back(g);
    g = help_gen_params{};
    g.kind = HELP_GEN_MCF; g.seed = 82; g.side = 8; g.commodities = 2; g.pairs = 4;
    gens.push_back(g);
    g = help_gen_params{};
    g.kind = HELP_GEN_STAIRCASE; g.seed = 83; g.periods = 6; g.block_rows = 60; g.block_cols = 180; g.block_k = 5;
    gens.push_back(g);
    g = help_gen_params{};
    g.kind = HELP_GEN_ZIPF; g.seed = 84; g.m = 500; g.nb = 1500; g.k = 6; g.zipf_s = 1.1;
    gens.push_back(g);
    for (size_t t = 0; t < gens.size(); ++t) {
        help_lp lp;
        CHECK(help_generate(&gens[t], &lp) == HELP_OK);
        char path[512];
        std::snprintf(path, sizeof path, "%s/inst%zu.bin", dir, t);
        CHECK(help_write_csc_bin(&lp, path) == HELP_OK);
        help_lp back;
        CHECK(help_read_csc_bin(path, &back) == HELP_OK);
        CHECK(back.nnz == lp.nnz && back.m == lp.m && back.n == lp.n && back.n_ub == lp.n_ub);
        help_lp_free(&back);
        std::snprintf(path, sizeof path, "%s/inst%zu.mps", dir, t);
        CHECK(help_write_mps(&lp, path, "SAN") == HELP_OK);
        symbolic_checks(lp, 0);
        symbolic_checks(lp, 1);
        if (gens[t].kind != HELP_GEN_ZIPF) {
            hsd_check(lp, HELP_HSD_EXACT);
            hsd_check(lp, 0);
        }
        help_lp_free(&lp);
    }
    // rejected inputs
    help_gen_params bad{};
    bad.kind = HELP_GEN_RANDOM; bad.m = 10; bad.nb = 5; bad.k = 20;
    help_lp lp;
    CHECK(help_generate(&bad, &lp) == HELP_ERR_ARG);
    CHECK(help_read_csc_bin("/nonexistent/file.bin", &lp) == HELP_ERR_ARG);
    std::printf("sanitized host checks: %d failure(s)\n", fails);
    return fails ? 1 : 0;
}
