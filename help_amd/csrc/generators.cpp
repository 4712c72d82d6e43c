// Deterministic LP instance generators and a fixed-format MPS writer.
//
// Every instance is in HELP's final-model layout (by default all columns
// x >= 0 with no upper bound, so n_lb = n, n_ub = n_fr = 0; the random LP
// can add bounded and free columns) and every value is rounded to
// six decimals, so that the text written by help_write_mps parses back
// (HELP MPSRead.cpp:150-170, atof at the fixed card positions of
// MPSRead.cpp:18) to exactly the doubles held here.
//
// Random numbers: splitmix64 seeds a xoshiro256** stream; normals use the
// Box-Muller transform.  The draw order is fixed per generator, so an
// instance is a pure function of (kind, sizes, seed).
#include "help_mi355x.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }           // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int32_t below(int32_t n) { return (int32_t)((next() >> 33) % (uint64_t)n); }
    double normal() {
        double u1 = 1.0 - uniform();   // (0,1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline double r6(double v) { return std::nearbyint(v * 1e6) / 1e6; }

// nonzero N(0,1) draw rounded to six decimals
inline double nz_normal(Rng& g) {
    for (;;) {
        double v = r6(g.normal());
        if (v != 0.0) return v;
    }
}

struct Builder {   // column-by-column CSC assembly
    int32_t m = 0;
    std::vector<int64_t> colptr{0};
    std::vector<int32_t> rowind;
    std::vector<double> val;
    std::vector<std::pair<int32_t, double>> tmp;
    void begin() { tmp.clear(); }
    void add(int32_t r, double v) { tmp.emplace_back(r, v); }
    void end() {
        std::sort(tmp.begin(), tmp.end(),
                  [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                      return a.first < b.first;
                  });
        for (auto& e : tmp) { rowind.push_back(e.first); val.push_back(e.second); }
        colptr.push_back((int64_t)rowind.size());
    }
    int32_t ncols() const { return (int32_t)colptr.size() - 1; }
};

// draw k distinct rows in [lo, lo+range)
void distinct_rows(Rng& g, int32_t lo, int32_t range, int32_t k, std::vector<int32_t>& out) {
    out.clear();
    while ((int32_t)out.size() < k) {
        int32_t r = lo + g.below(range);
        if (std::find(out.begin(), out.end(), r) == out.end()) out.push_back(r);
    }
}

int finish(Builder& B, std::vector<double>& b, std::vector<double>& c, help_lp* lp) {
    lp->m = B.m;
    lp->n = B.ncols();
    lp->nnz = (int64_t)B.rowind.size();
    lp->n_ub = 0; lp->n_lb = lp->n; lp->n_fr = 0;
    lp->c_intercept = 0.0;
    lp->colptr = (int64_t*)std::malloc(sizeof(int64_t) * (lp->n + 1));
    lp->rowind = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(1, lp->nnz));
    lp->val = (double*)std::malloc(sizeof(double) * std::max<int64_t>(1, lp->nnz));
    lp->b = (double*)std::malloc(sizeof(double) * std::max(1, lp->m));
    lp->c = (double*)std::malloc(sizeof(double) * std::max(1, lp->n));
    lp->ub = nullptr;
    if (!lp->colptr || !lp->rowind || !lp->val || !lp->b || !lp->c) return HELP_ERR_NOMEM;
    std::memcpy(lp->colptr, B.colptr.data(), sizeof(int64_t) * (lp->n + 1));
    std::memcpy(lp->rowind, B.rowind.data(), sizeof(int32_t) * lp->nnz);
    std::memcpy(lp->val, B.val.data(), sizeof(double) * lp->nnz);
    std::memcpy(lp->b, b.data(), sizeof(double) * lp->m);
    std::memcpy(lp->c, c.data(), sizeof(double) * lp->n);
    return HELP_OK;
}

// b = A x0, c = A^T y0 + s0 with x0, s0 ~ U(0.5,1.5), y0 ~ N(0,1): strictly
// primal and dual feasible, hence bounded.
void feasible_rhs_cost(Rng& g, const Builder& B, std::vector<double>& b, std::vector<double>& c) {
    int32_t n = B.ncols(), m = B.m;
    std::vector<double> x0(n), y0(m), s0(n);
    for (int32_t j = 0; j < n; ++j) x0[j] = g.uniform(0.5, 1.5);
    for (int32_t i = 0; i < m; ++i) y0[i] = g.normal();
    for (int32_t j = 0; j < n; ++j) s0[j] = g.uniform(0.5, 1.5);
    b.assign(m, 0.0);
    c.assign(n, 0.0);
    for (int32_t j = 0; j < n; ++j) {
        double cj = 0.0;
        for (int64_t p = B.colptr[j]; p < B.colptr[j + 1]; ++p) {
            b[B.rowind[p]] += B.val[p] * x0[j];
            cj += B.val[p] * y0[B.rowind[p]];
        }
        c[j] = r6(cj + s0[j]);
    }
    for (int32_t i = 0; i < m; ++i) b[i] = r6(b[i]);
}

void add_identity(Builder& B) {
    for (int32_t i = 0; i < B.m; ++i) { B.begin(); B.add(i, 1.0); B.end(); }
}

// Bounded / free variant of the random LP, in HELP's final column layout
// (Presolve.cpp:1101-1170): the first nub columns of B carry upper bounds,
// the rest of B and the identity a lower bound only, and nfr extra columns
// (k nnz each) after the identity are free.  Interior point: x0 ~ U(0.5,1.5)
// with u = x0 + U(0.5,1.5) on bounded columns, x0 ~ N(0,1) on free ones;
// y0 ~ N(0,1); reduced costs s0 ~ U(0.5,1.5), zero on free columns (their
// dual row is an equality).
int finish_bounded(Rng& g, Builder& B, int32_t nub, int32_t nfr, help_lp* lp) {
    const int32_t n = B.ncols(), m = B.m;
    std::vector<double> x0(n), y0(m), s0(n, 0.0), u(nub);
    for (int32_t j = 0; j < n - nfr; ++j) x0[j] = g.uniform(0.5, 1.5);
    for (int32_t j = n - nfr; j < n; ++j) x0[j] = g.normal();
    for (int32_t j = 0; j < nub; ++j) u[j] = r6(x0[j] + g.uniform(0.5, 1.5));
    for (int32_t i = 0; i < m; ++i) y0[i] = g.normal();
    for (int32_t j = 0; j < n - nfr; ++j) s0[j] = g.uniform(0.5, 1.5);
    std::vector<double> b(m, 0.0), c(n, 0.0);
    for (int32_t j = 0; j < n; ++j) {
        double cj = 0.0;
        for (int64_t p = B.colptr[j]; p < B.colptr[j + 1]; ++p) {
            b[B.rowind[p]] += B.val[p] * x0[j];
            cj += B.val[p] * y0[B.rowind[p]];
        }
        c[j] = r6(cj + s0[j]);
    }
    for (int32_t i = 0; i < m; ++i) b[i] = r6(b[i]);
    int st = finish(B, b, c, lp);
    if (st != HELP_OK) return st;
    lp->n_ub = nub;
    lp->n_fr = nfr;
    lp->n_lb = n - nfr;
    if (nub > 0) {
        lp->ub = (double*)std::malloc(sizeof(double) * nub);
        if (!lp->ub) return HELP_ERR_NOMEM;
        std::memcpy(lp->ub, u.data(), sizeof(double) * nub);
    }
    return HELP_OK;
}

int gen_random(const help_gen_params* p, help_lp* lp) {
    if (p->m <= 0 || p->nb < 0 || p->k <= 0 || p->k > p->m) return HELP_ERR_ARG;
    if (p->ub_cols < 0 || p->ub_cols > p->nb || p->free_cols < 0) return HELP_ERR_ARG;
    Rng g(p->seed);
    Builder B; B.m = p->m;
    std::vector<int32_t> rows;
    for (int32_t j = 0; j < p->nb; ++j) {
        distinct_rows(g, 0, p->m, p->k, rows);
        B.begin();
        for (int32_t r : rows) B.add(r, nz_normal(g));
        B.end();
    }
    add_identity(B);
    if (p->ub_cols > 0 || p->free_cols > 0) {
        for (int32_t j = 0; j < p->free_cols; ++j) {
            distinct_rows(g, 0, p->m, p->k, rows);
            B.begin();
            for (int32_t r : rows) B.add(r, nz_normal(g));
            B.end();
        }
        return finish_bounded(g, B, p->ub_cols, p->free_cols, lp);
    }
    std::vector<double> b, c;
    feasible_rhs_cost(g, B, b, c);
    return finish(B, b, c, lp);
}

int gen_zipf(const help_gen_params* p, help_lp* lp) {
    if (p->m <= 0 || p->nb < 0 || p->k <= 0 || p->k > p->m || !(p->zipf_s > 0)) return HELP_ERR_ARG;
    Rng g(p->seed);
    Builder B; B.m = p->m;
    // inverse-CDF table of P(rank r) ~ r^-s, r = 1..m; rank r -> row r-1
    std::vector<double> cdf(p->m);
    double acc = 0.0;
    for (int32_t r = 0; r < p->m; ++r) { acc += std::pow((double)(r + 1), -p->zipf_s); cdf[r] = acc; }
    for (auto& v : cdf) v /= acc;
    cdf[p->m - 1] = 1.0;
    std::vector<int32_t> rows;
    for (int32_t j = 0; j < p->nb; ++j) {
        rows.clear();
        while ((int32_t)rows.size() < p->k) {
            double u = g.uniform();
            int32_t r = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            if (r >= p->m) r = p->m - 1;
            if (std::find(rows.begin(), rows.end(), r) == rows.end()) rows.push_back(r);
        }
        B.begin();
        for (int32_t r : rows) B.add(r, nz_normal(g));
        B.end();
    }
    add_identity(B);
    std::vector<double> b, c;
    feasible_rhs_cost(g, B, b, c);
    return finish(B, b, c, lp);
}

int gen_staircase(const help_gen_params* p, help_lp* lp) {
    int32_t T = p->periods, R = p->block_rows, C = p->block_cols, k = p->block_k;
    if (T <= 0 || R <= 0 || C <= 0 || k <= 0 || k > R) return HELP_ERR_ARG;
    Rng g(p->seed);
    Builder B; B.m = T * R;
    std::vector<int32_t> rows;
    for (int32_t t = 0; t < T; ++t)
        for (int32_t j = 0; j < C; ++j) {
            distinct_rows(g, t * R, R, k, rows);
            B.begin();
            for (int32_t r : rows) B.add(r, nz_normal(g));
            if (t + 1 < T) {
                int32_t r = (t + 1) * R + g.below(R);
                B.add(r, nz_normal(g));
            }
            B.end();
        }
    add_identity(B);
    std::vector<double> b, c;
    feasible_rhs_cost(g, B, b, c);
    return finish(B, b, c, lp);
}

// Block-angular multicommodity flow on a side x side 4-connected grid.
// Rows: per commodity, node-arc conservation for nodes 0..N-2 (the last node
// row is dropped for full row rank), then one capacity row per arc.
// Columns: x_{k,a} for each commodity k and arc a, then one slack per arc.
int gen_mcf(const help_gen_params* p, help_lp* lp) {
    int32_t G = p->side, K = p->commodities, P = p->pairs;
    if (G < 2 || K <= 0 || P < 0) return HELP_ERR_ARG;
    Rng g(p->seed);
    int32_t N = G * G;
    std::vector<int32_t> tail, head;
    const int dr[4] = {0, 1, 0, -1}, dc[4] = {1, 0, -1, 0};
    for (int32_t v = 0; v < N; ++v) {
        int32_t r = v / G, c = v % G;
        for (int d = 0; d < 4; ++d) {
            int32_t rr = r + dr[d], cc = c + dc[d];
            if (rr < 0 || rr >= G || cc < 0 || cc >= G) continue;
            tail.push_back(v);
            head.push_back(rr * G + cc);
        }
    }
    int32_t Na = (int32_t)tail.size();
    std::vector<double> cost(Na), cap(Na);
    for (int32_t a = 0; a < Na; ++a) cost[a] = r6(g.uniform(1.0, 10.0));
    for (int32_t a = 0; a < Na; ++a) cap[a] = r6(g.uniform(5.0, 50.0));
    int32_t ncons = N - 1;
    Builder B; B.m = K * ncons + Na;
    std::vector<double> b(B.m, 0.0), c;
    for (int32_t k = 0; k < K; ++k)
        for (int32_t q = 0; q < P; ++q) {
            int32_t s = g.below(N), t;
            do { t = g.below(N); } while (t == s);
            double dem = r6(g.uniform(1.0, 10.0));
            if (s < ncons) b[(size_t)k * ncons + s] += dem;
            if (t < ncons) b[(size_t)k * ncons + t] -= dem;
        }
    for (int32_t i = 0; i < K * ncons; ++i) b[i] = r6(b[i]);
    for (int32_t a = 0; a < Na; ++a) b[(size_t)K * ncons + a] = cap[a];
    for (int32_t k = 0; k < K; ++k)
        for (int32_t a = 0; a < Na; ++a) {
            B.begin();
            if (tail[a] < ncons) B.add(k * ncons + tail[a], 1.0);
            if (head[a] < ncons) B.add(k * ncons + head[a], -1.0);
            B.add(K * ncons + a, 1.0);
            B.end();
            c.push_back(cost[a]);
        }
    for (int32_t a = 0; a < Na; ++a) {
        B.begin(); B.add(K * ncons + a, 1.0); B.end();
        c.push_back(0.0);
    }
    return finish(B, b, c, lp);
}

}  // namespace

extern "C" int help_generate(const help_gen_params* p, help_lp* lp) {
    if (!p || !lp) return HELP_ERR_ARG;
    std::memset(lp, 0, sizeof(*lp));
    switch (p->kind) {
        case HELP_GEN_RANDOM: return gen_random(p, lp);
        case HELP_GEN_MCF: return gen_mcf(p, lp);
        case HELP_GEN_STAIRCASE: return gen_staircase(p, lp);
        case HELP_GEN_ZIPF: return gen_zipf(p, lp);
        default: return HELP_ERR_ARG;
    }
}

extern "C" void help_lp_free(help_lp* lp) {
    if (!lp) return;
    std::free(lp->colptr); std::free(lp->rowind); std::free(lp->val);
    std::free(lp->b); std::free(lp->c); std::free(lp->ub);
    std::memset(lp, 0, sizeof(*lp));
}

// Fixed-format MPS: names at card positions 4 and 14, value at 24
// (MPSRead.cpp:18), one matrix entry per line, names of at most 8 chars
// (MPSRead.cpp:26-45 hashes the first 8).  Zero costs and zero right-hand
// sides are omitted; MPSRead leaves them at 0.
extern "C" int help_write_mps(const help_lp* lp, const char* path, const char* name) {
    if (!lp || !path) return HELP_ERR_ARG;
    if (lp->m > 9999999 || lp->n > 9999999 || lp->n_ub < 0 || lp->n_fr < 0 ||
        lp->n_ub + lp->n_fr > lp->n || (lp->n_ub > 0 && !lp->ub))
        return HELP_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return HELP_ERR_ARG;
    std::fprintf(f, "NAME          %s\n", name ? name : "HELPGEN");
    std::fprintf(f, "ROWS\n N  COST\n");
    for (int32_t i = 0; i < lp->m; ++i) std::fprintf(f, " E  R%07d\n", i);
    std::fprintf(f, "COLUMNS\n");
    for (int32_t j = 0; j < lp->n; ++j) {
        if (lp->c[j] != 0.0) std::fprintf(f, "    C%07d  COST      %.6f\n", j, lp->c[j]);
        for (int64_t p = lp->colptr[j]; p < lp->colptr[j + 1]; ++p)
            std::fprintf(f, "    C%07d  R%07d  %.6f\n", j, lp->rowind[p], lp->val[p]);
    }
    std::fprintf(f, "RHS\n");
    for (int32_t i = 0; i < lp->m; ++i)
        if (lp->b[i] != 0.0) std::fprintf(f, "    RHS       R%07d  %.6f\n", i, lp->b[i]);
    // bounds at card positions 1 (type), 4 (set), 14 (column), 24 (value)
    // as MPSRead.cpp:286-345 reads them; LB stays at its default 0
    if (lp->n_ub > 0 || lp->n_fr > 0) {
        std::fprintf(f, "BOUNDS\n");
        for (int32_t j = 0; j < lp->n_ub; ++j) std::fprintf(f, " UP BND       C%07d  %.6f\n", j, lp->ub[j]);
        for (int32_t j = lp->n - lp->n_fr; j < lp->n; ++j) std::fprintf(f, " FR BND       C%07d\n", j);
    }
    std::fprintf(f, "ENDATA\n");
    std::fclose(f);
    return HELP_OK;
}

// ---------------------------------------------------------------------------
// Binary CSC instance file, for LPs beyond the fixed-MPS limits (HELP's
// MAX_ROWS / MAX_COLS / MAX_ELEMENTS of LP.h:61-63, e.g. m = 1e6, n = 4e6).
// Little-endian layout:
//   char[8]  "HELPCSC1"
//   int32    m, n, n_ub, n_lb, n_fr, reserved (0)
//   int64    nnz
//   double   c_intercept
//   int64    colptr[n + 1]
//   int32    rowind[nnz]
//   double   val[nnz], b[m], c[n], ub[n_ub]
//   uint64   FNV-1a hash of every byte above
// ---------------------------------------------------------------------------
namespace {
constexpr char kCscMagic[8] = {'H', 'E', 'L', 'P', 'C', 'S', 'C', '1'};

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 0x100000001b3ull; }
    }
};

bool put(FILE* f, Fnv& fnv, const void* p, size_t n) {
    if (n == 0) return true;
    fnv.add(p, n);
    return std::fwrite(p, 1, n, f) == n;
}

bool get(FILE* f, Fnv& fnv, void* p, size_t n) {
    if (n == 0) return true;
    if (std::fread(p, 1, n, f) != n) return false;
    fnv.add(p, n);
    return true;
}
}  // namespace

extern "C" int help_write_csc_bin(const help_lp* lp, const char* path) {
    if (!lp || !path || lp->m < 0 || lp->n < 0 || lp->nnz < 0 || lp->n_ub < 0 || lp->n_fr < 0 ||
        (lp->n_ub > 0 && !lp->ub))
        return HELP_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return HELP_ERR_ARG;
    Fnv fnv;
    const int32_t hdr[6] = {lp->m, lp->n, lp->n_ub, lp->n_lb, lp->n_fr, 0};
    bool ok = put(f, fnv, kCscMagic, 8) && put(f, fnv, hdr, sizeof hdr) && put(f, fnv, &lp->nnz, 8) &&
              put(f, fnv, &lp->c_intercept, 8) && put(f, fnv, lp->colptr, sizeof(int64_t) * (lp->n + 1)) &&
              put(f, fnv, lp->rowind, sizeof(int32_t) * lp->nnz) && put(f, fnv, lp->val, sizeof(double) * lp->nnz) &&
              put(f, fnv, lp->b, sizeof(double) * lp->m) && put(f, fnv, lp->c, sizeof(double) * lp->n) &&
              put(f, fnv, lp->ub, sizeof(double) * lp->n_ub);
    const uint64_t h = fnv.h;
    ok = ok && std::fwrite(&h, 1, 8, f) == 8;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HELP_OK : HELP_ERR_ARG;
}

extern "C" int help_read_csc_bin(const char* path, help_lp* lp) {
    if (!path || !lp) return HELP_ERR_ARG;
    std::memset(lp, 0, sizeof(*lp));
    FILE* f = std::fopen(path, "rb");
    if (!f) return HELP_ERR_ARG;
    Fnv fnv;
    char magic[8];
    int32_t hdr[6];
    int64_t nnz = 0;
    double icpt = 0;
    int st = HELP_ERR_ARG;
    if (get(f, fnv, magic, 8) && std::memcmp(magic, kCscMagic, 8) == 0 && get(f, fnv, hdr, sizeof hdr) &&
        get(f, fnv, &nnz, 8) && get(f, fnv, &icpt, 8)) {
        const int32_t m = hdr[0], n = hdr[1], nub = hdr[2], nlb = hdr[3], nfr = hdr[4];
        if (m >= 0 && n >= 0 && nnz >= 0 && nnz <= INT32_MAX && nub >= 0 && nfr >= 0 && nub + nfr <= n &&
            nlb == n - nfr) {
            lp->m = m; lp->n = n; lp->nnz = nnz; lp->n_ub = nub; lp->n_lb = nlb; lp->n_fr = nfr;
            lp->c_intercept = icpt;
            lp->colptr = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
            lp->rowind = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(1, nnz));
            lp->val = (double*)std::malloc(sizeof(double) * std::max<int64_t>(1, nnz));
            lp->b = (double*)std::malloc(sizeof(double) * std::max(1, m));
            lp->c = (double*)std::malloc(sizeof(double) * std::max(1, n));
            lp->ub = nub > 0 ? (double*)std::malloc(sizeof(double) * nub) : nullptr;
            uint64_t h = 0;
            if (!lp->colptr || !lp->rowind || !lp->val || !lp->b || !lp->c || (nub > 0 && !lp->ub)) {
                st = HELP_ERR_NOMEM;
            } else if (get(f, fnv, lp->colptr, sizeof(int64_t) * (n + 1)) &&
                       get(f, fnv, lp->rowind, sizeof(int32_t) * nnz) && get(f, fnv, lp->val, sizeof(double) * nnz) &&
                       get(f, fnv, lp->b, sizeof(double) * m) && get(f, fnv, lp->c, sizeof(double) * n) &&
                       get(f, fnv, lp->ub, sizeof(double) * nub) && std::fread(&h, 1, 8, f) == 8 && h == fnv.h) {
                // structural checks: monotone column pointers, rows in range
                bool good = lp->colptr[0] == 0 && lp->colptr[n] == nnz;
                for (int32_t j = 0; good && j < n; ++j) good = lp->colptr[j] <= lp->colptr[j + 1];
                for (int64_t p = 0; good && p < nnz; ++p) good = lp->rowind[p] >= 0 && lp->rowind[p] < m;
                if (good) st = HELP_OK;
            }
        }
    }
    std::fclose(f);
    if (st != HELP_OK) help_lp_free(lp);
    return st;
}
