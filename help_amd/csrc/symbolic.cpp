// Host symbolic analysis: ADA^T pattern, AMD ordering, etree, postorder,
// column counts, relaxed supernodes, relative maps, level schedule.
#include "symbolic.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <memory>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <numeric>
#include <stdexcept>

namespace help {

CsrMatrix csc_to_csr(const CscMatrix& A) {
    CsrMatrix R;
    R.m = A.m; R.n = A.n;
    int64_t nnz = A.colptr[A.n];
    R.rowptr.assign(A.m + 1, 0);
    for (int64_t p = 0; p < nnz; ++p) R.rowptr[A.rowind[p] + 1]++;
    for (int32_t i = 0; i < A.m; ++i) R.rowptr[i + 1] += R.rowptr[i];
    R.colind.resize(nnz);
    R.cscpos.resize(nnz);
    std::vector<int64_t> fill(R.rowptr.begin(), R.rowptr.end() - 1);
    for (int32_t j = 0; j < A.n; ++j)          // ascending j -> columns sorted per row
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
            int64_t t = fill[A.rowind[p]]++;
            R.colind[t] = j;
            R.cscpos[t] = p;
        }
    return R;
}

// HELP Helper.cpp:273-329 (ADAt_Allocate, linker form) gives the slot order;
// Helper.cpp:350-377 (ADAt_Calc) gives the order of the contributions.
AdatPattern adat_pattern(const CscMatrix& A, const CsrMatrix& R, bool with_pairs) {
    // Rows are independent: each thread keeps its own marker arrays, and
    // every row's slots and pairs come out in the same order whatever the
    // thread count.
    AdatPattern P;
    const int32_t m = A.m;
    P.rowptr.assign(m + 1, 0);
    // pass 1: slots per row
#pragma omp parallel
    {
        std::vector<int32_t> flag(m, -1);
#pragma omp for schedule(dynamic, 512)
        for (int32_t i = 0; i < m; ++i) {
            int64_t cnt = 0;
            for (int64_t t = R.rowptr[i]; t < R.rowptr[i + 1]; ++t) {
                int32_t j = R.colind[t];
                for (int64_t q = R.cscpos[t]; q < A.colptr[j + 1]; ++q) {
                    int32_t k = A.rowind[q];
                    if (flag[k] != i) { flag[k] = i; ++cnt; }
                }
            }
            P.rowptr[i + 1] = cnt;
        }
    }
    for (int32_t i = 0; i < m; ++i) P.rowptr[i + 1] += P.rowptr[i];
    const int64_t nup = P.rowptr[m];
    P.colind.resize(nup);
    std::vector<int64_t> npair;
    if (with_pairs) npair.assign(nup, 0);
    // pass 2: slots in discovery order (and pair counts)
#pragma omp parallel
    {
        std::vector<int32_t> flag(m, -1), slot(m, 0);
#pragma omp for schedule(dynamic, 512)
        for (int32_t i = 0; i < m; ++i) {
            int64_t s = P.rowptr[i];
            for (int64_t t = R.rowptr[i]; t < R.rowptr[i + 1]; ++t) {
                int32_t j = R.colind[t];
                for (int64_t q = R.cscpos[t]; q < A.colptr[j + 1]; ++q) {
                    int32_t k = A.rowind[q];
                    if (flag[k] != i) { flag[k] = i; slot[k] = (int32_t)(s - P.rowptr[i]); P.colind[s++] = k; }
                    if (with_pairs) npair[P.rowptr[i] + slot[k]]++;
                }
            }
        }
    }
    if (!with_pairs) return P;
    P.pairptr.assign(nup + 1, 0);
    for (int64_t s = 0; s < nup; ++s) P.pairptr[s + 1] = P.pairptr[s] + npair[s];
    int64_t np = P.pairptr[nup];
    if (np > INT32_MAX) throw std::runtime_error("too many ADA^T contribution pairs");
    P.pair_p.resize(np);
    P.pair_q.resize(np);
    // pass 3: each slot's (p, q) products in the order of the row's entries
    // (a slot's fill pointer belongs to its row alone)
    std::vector<int64_t> fillp(P.pairptr.begin(), P.pairptr.end() - 1);
#pragma omp parallel
    {
        std::vector<int32_t> slot(m, 0);
#pragma omp for schedule(dynamic, 512)
        for (int32_t i = 0; i < m; ++i) {
            const int64_t base = P.rowptr[i];
            for (int64_t s = base; s < P.rowptr[i + 1]; ++s) slot[P.colind[s]] = (int32_t)(s - base);
            for (int64_t t = R.rowptr[i]; t < R.rowptr[i + 1]; ++t) {
                int32_t j = R.colind[t];
                int64_t p = R.cscpos[t];
                for (int64_t q = p; q < A.colptr[j + 1]; ++q) {
                    int64_t f = fillp[base + slot[A.rowind[q]]]++;
                    P.pair_p[f] = (int32_t)p;
                    P.pair_q[f] = (int32_t)q;
                }
            }
        }
    }
    return P;
}

void adat_adjacency(const CscMatrix& A, const CsrMatrix& R, std::vector<int64_t>& adjptr,
                    std::vector<int32_t>& adj) {
    // row i of the symmetric pattern (without the diagonal) = every row that
    // shares a column of A with row i: each row is discovered on its own
    // (thread-local markers, no atomics), counted, then filled and sorted
    const int32_t m = A.m;
    adjptr.assign(m + 1, 0);
#pragma omp parallel
    {
        std::vector<int32_t> flag(m, -1);
#pragma omp for schedule(dynamic, 512)
        for (int32_t i = 0; i < m; ++i) {
            int64_t cnt = 0;
            flag[i] = i;
            for (int64_t t = R.rowptr[i]; t < R.rowptr[i + 1]; ++t) {
                const int32_t j = R.colind[t];
                for (int64_t q = A.colptr[j]; q < A.colptr[j + 1]; ++q) {
                    const int32_t k = A.rowind[q];
                    if (flag[k] != i) { flag[k] = i; ++cnt; }
                }
            }
            adjptr[i + 1] = cnt;
        }
    }
    for (int32_t i = 0; i < m; ++i) adjptr[i + 1] += adjptr[i];
    adj.resize(adjptr[m]);
#pragma omp parallel
    {
        std::vector<int32_t> flag(m, -1);
#pragma omp for schedule(dynamic, 512)
        for (int32_t i = 0; i < m; ++i) {
            int64_t o = adjptr[i];
            flag[i] = i;
            for (int64_t t = R.rowptr[i]; t < R.rowptr[i + 1]; ++t) {
                const int32_t j = R.colind[t];
                for (int64_t q = A.colptr[j]; q < A.colptr[j + 1]; ++q) {
                    const int32_t k = A.rowind[q];
                    if (flag[k] != i) { flag[k] = i; adj[o++] = k; }
                }
            }
            std::sort(adj.begin() + adjptr[i], adj.begin() + adjptr[i + 1]);
        }
    }
}

void permute_graph(int32_t m, const std::vector<int64_t>& adjptr, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& perm, const std::vector<int32_t>& iperm,
                   std::vector<int64_t>& pp, std::vector<int32_t>& pj, int half) {
    pp.assign(m + 1, 0);
    auto keep = [&](int32_t i, int32_t k) { return half == 0 || (half < 0 ? i < k : i > k); };
    if (half == 0) {
        for (int32_t k = 0; k < m; ++k) pp[k + 1] = pp[k] + (adjptr[perm[k] + 1] - adjptr[perm[k]]);
    } else {
#pragma omp parallel for schedule(dynamic, 512)
        for (int32_t k = 0; k < m; ++k) {
            const int32_t old = perm[k];
            int64_t cnt = 0;
            for (int64_t t = adjptr[old]; t < adjptr[old + 1]; ++t) cnt += keep(iperm[adj[t]], k) ? 1 : 0;
            pp[k + 1] = cnt;
        }
        for (int32_t k = 0; k < m; ++k) pp[k + 1] += pp[k];
    }
    pj.resize(pp[m]);
#pragma omp parallel for schedule(dynamic, 512)
    for (int32_t k = 0; k < m; ++k) {
        const int32_t old = perm[k];
        int64_t o = pp[k];
        for (int64_t t = adjptr[old]; t < adjptr[old + 1]; ++t) {
            const int32_t i = iperm[adj[t]];
            if (keep(i, k)) pj[o++] = i;
        }
    }
}

// ---------------------------------------------------------------------------
// Approximate minimum degree on the quotient graph (Amestoy, Davis & Duff,
// SIAM J. Matrix Anal. Appl. 17(4), 1996): element absorption, aggressive
// absorption, approximate external degrees, indistinguishable-variable
// (supervariable) detection by hashing, mass elimination and postponement
// of dense rows (degree > max(16, alpha*sqrt(n))).
// ---------------------------------------------------------------------------
std::vector<int32_t> amd_order(int32_t n, const std::vector<int64_t>& adjptr,
                               const std::vector<int32_t>& adj, double dense_alpha) {
    enum : int8_t { VAR = 0, ELEM = 1, DEAD = 2 };
    std::vector<int32_t> order;
    order.reserve(n);
    if (n == 0) return order;
    std::vector<int32_t> nv(n, 1), sv_next(n, -1), sv_tail(n);
    std::vector<int8_t> status(n, VAR);
    std::vector<int64_t> deg(n, 0), edeg(n, 0), extdeg(n, 0), w(n, 0);
    std::vector<uint64_t> hsh(n, 0);
    std::vector<std::vector<int32_t>> E(n), Av(n), Le(n);
    std::iota(sv_tail.begin(), sv_tail.end(), 0);

    int64_t dense = std::max<int64_t>(16, (int64_t)(dense_alpha * std::sqrt((double)n)));
    std::vector<int32_t> dense_nodes;
    for (int32_t i = 0; i < n; ++i)
        if (adjptr[i + 1] - adjptr[i] > dense) { status[i] = DEAD; dense_nodes.push_back(i); }
    int64_t nlive = n - (int64_t)dense_nodes.size();

    std::vector<int32_t> head(n + 1, -1), nxt(n, -1), prv(n, -1);
    auto bucket_insert = [&](int32_t i, int64_t d) {
        nxt[i] = head[d]; prv[i] = -1;
        if (head[d] >= 0) prv[head[d]] = i;
        head[d] = i;
    };
    auto bucket_remove = [&](int32_t i) {
        int64_t d = deg[i];
        if (prv[i] >= 0) nxt[prv[i]] = nxt[i]; else head[d] = nxt[i];
        if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
        nxt[i] = prv[i] = -1;
    };
    for (int32_t i = 0; i < n; ++i) {
        if (status[i] != VAR) continue;
        for (int64_t t = adjptr[i]; t < adjptr[i + 1]; ++t)
            if (status[adj[t]] == VAR) Av[i].push_back(adj[t]);
        deg[i] = (int64_t)Av[i].size();
    }
    for (int32_t i = n - 1; i >= 0; --i)
        if (status[i] == VAR) bucket_insert(i, deg[i]);

    // vm[j]: -1 unless j is a live variable (status VAR, nv > 0); else its
    // latest stamp (membership of the current L_p, or of the pattern being
    // compared in supervariable detection): one load answers both questions.
    // em[e]: the same stamps for elements.
    std::vector<int32_t> vm(n, 0), em(n, 0);
    for (int32_t i = 0; i < n; ++i)
        if (status[i] != VAR) vm[i] = -1;
    int32_t stamp = 0;
    int64_t wflg = 1;
    int64_t nel = 0, mindeg = 0;
    auto emit = [&](int32_t v) {   // order a whole supervariable
        for (int32_t j = v; j >= 0; j = sv_next[j]) order.push_back(j);
    };
    auto next_stamp = [&]() {
        if (stamp == INT32_MAX) {
            for (int32_t i = 0; i < n; ++i) { if (vm[i] > 0) vm[i] = 0; em[i] = 0; }
            stamp = 0;
        }
        return ++stamp;
    };
    std::vector<int32_t> Lp_live;
    std::vector<std::pair<uint64_t, int32_t>> hl;

    while (nel < nlive) {
        while (head[mindeg] < 0) ++mindeg;
        int32_t p = head[mindeg];
        bucket_remove(p);
        // --- new element p: L_p = (A_p U L_e for e in E_p) \ {p} ---
        int32_t st = next_stamp();
        vm[p] = -1;   // p becomes an element
        std::vector<int32_t>& Lp = Le[p];
        Lp.clear();
        int64_t degme = 0;
        for (int32_t j : Av[p])
            if (vm[j] >= 0 && vm[j] != st) { vm[j] = st; Lp.push_back(j); degme += nv[j]; }
        for (int32_t e : E[p]) {
            if (status[e] != ELEM) continue;
            for (int32_t j : Le[e])
                if (vm[j] >= 0 && vm[j] != st) { vm[j] = st; Lp.push_back(j); degme += nv[j]; }
            status[e] = DEAD;
            std::vector<int32_t>().swap(Le[e]);
        }
        std::vector<int32_t>().swap(Av[p]);
        std::vector<int32_t>().swap(E[p]);
        status[p] = ELEM;
        nel += nv[p];
        emit(p);
        for (int32_t i : Lp) bucket_remove(i);
        // --- |L_e \ L_p| for every element adjacent to L_p ---
        wflg += (int64_t)n + 2;
        const size_t nLp = Lp.size();
        // the lists of L_p's variables are scattered over the heap and the
        // two scans below are bound by the latency of reaching them: fetch
        // the list headers four variables ahead and their first lines two
        // ahead
        auto prefetch_lists = [&](size_t x) {
            if (x + 4 < nLp) { __builtin_prefetch(&E[Lp[x + 4]]); __builtin_prefetch(&Av[Lp[x + 4]]); }
            if (x + 2 < nLp) {
                const int32_t* pe = E[Lp[x + 2]].data();
                const int32_t* pa = Av[Lp[x + 2]].data();
                __builtin_prefetch(pe); __builtin_prefetch(pa); __builtin_prefetch(pa + 16);
            }
        };
        for (size_t x = 0; x < nLp; ++x) {
            prefetch_lists(x);
            const int32_t i = Lp[x];
            for (int32_t e : E[i]) {
                if (status[e] != ELEM) continue;
                if (w[e] < wflg) w[e] = wflg + edeg[e];
                w[e] -= nv[i];
            }
        }
        // --- prune lists, external degrees, mass elimination, hashes ---
        Lp_live.clear();
        for (size_t x = 0; x < nLp; ++x) {
            prefetch_lists(x);
            const int32_t i = Lp[x];
            int64_t ext = 0;
            uint64_t h = 0;
            size_t k = 0;
            for (int32_t e : E[i]) {
                if (status[e] != ELEM) continue;
                int64_t we = w[e] - wflg;
                if (we <= 0) {   // aggressive absorption: L_e is inside L_p
                    status[e] = DEAD;
                    std::vector<int32_t>().swap(Le[e]);
                    continue;
                }
                ext += we;
                h += (uint64_t)e;
                E[i][k++] = e;
            }
            E[i].resize(k);
            k = 0;
            for (int32_t j : Av[i]) {
                const int32_t vj = vm[j];
                if (vj < 0 || vj == st) continue;
                ext += nv[j];
                h += (uint64_t)j;
                Av[i][k++] = j;
            }
            Av[i].resize(k);
            if (E[i].empty() && Av[i].empty()) {   // mass elimination with p
                status[i] = DEAD;
                vm[i] = -1;
                nel += nv[i];
                degme -= nv[i];
                emit(i);
                std::vector<int32_t>().swap(E[i]);
                std::vector<int32_t>().swap(Av[i]);
                continue;
            }
            E[i].push_back(p);
            h += (uint64_t)p;
            extdeg[i] = ext;
            hsh[i] = h;
            Lp_live.push_back(i);
        }
        // --- supervariable detection among L_p ---
        hl.clear();
        for (int32_t i : Lp_live) hl.emplace_back(hsh[i], i);
        std::sort(hl.begin(), hl.end());
        for (size_t a = 0; a < hl.size();) {
            size_t b = a;
            while (b < hl.size() && hl[b].first == hl[a].first) ++b;
            for (size_t x = a; x < b; ++x) {
                int32_t i = hl[x].second;
                if (nv[i] == 0) continue;
                bool marked = false;
                int32_t st2 = 0;
                for (size_t y = x + 1; y < b; ++y) {
                    int32_t j = hl[y].second;
                    if (nv[j] == 0) continue;
                    if (E[i].size() != E[j].size() || Av[i].size() != Av[j].size()) continue;
                    if (!marked) {
                        st2 = next_stamp();
                        for (int32_t e : E[i]) em[e] = st2;
                        for (int32_t v : Av[i]) vm[v] = st2;
                        marked = true;
                    }
                    bool same = true;
                    for (int32_t e : E[j]) if (em[e] != st2) { same = false; break; }
                    if (same) for (int32_t v : Av[j]) if (vm[v] != st2) { same = false; break; }
                    if (!same) continue;
                    nv[i] += nv[j];
                    nv[j] = 0;
                    status[j] = DEAD;
                    vm[j] = -1;
                    sv_next[sv_tail[i]] = j;
                    sv_tail[i] = sv_tail[j];
                    std::vector<int32_t>().swap(E[j]);
                    std::vector<int32_t>().swap(Av[j]);
                }
            }
            a = b;
        }
        // --- finalize degrees and L_p ---
        size_t k = 0;
        for (int32_t i : Lp_live) {
            if (nv[i] == 0) continue;
            int64_t d = std::min(deg[i] + degme - nv[i], extdeg[i] + degme - nv[i]);
            d = std::min(d, nlive - nel - nv[i]);
            if (d < 0) d = 0;
            deg[i] = d;
            bucket_insert(i, d);
            if (d < mindeg) mindeg = d;
            Lp[k++] = i;
        }
        Lp.resize(k);
        edeg[p] = degme;
    }
    for (int32_t i : dense_nodes) order.push_back(i);
    if ((int32_t)order.size() != n) throw std::runtime_error("amd_order: incomplete ordering");
    return order;
}

// ---------------------------------------------------------------------------
// Nested dissection by level-structure bisection (George & Liu, "Computer
// Solution of Large Sparse Positive Definite Systems", 1981, ch. 8): a BFS
// level structure rooted at a pseudo-peripheral node, a separator level near
// the middle chosen for small size, trimmed to the vertices that touch the
// far side, the two sides dissected recursively and the separator numbered
// after them.  Subgraphs at or below `leaf` vertices, and subgraphs whose
// best separator exceeds half of them, are ordered by approximate minimum
// degree.  Disconnected subgraphs are dissected per component.  Entirely
// deterministic.
// ---------------------------------------------------------------------------
namespace {

using NdJobs = std::vector<std::vector<int32_t>>;

struct NdWork {
    const std::vector<int64_t>& ptr;
    const std::vector<int32_t>& adj;
    // membership / visit marks: a subgraph's vertices carry its tag; the two
    // halves of a bisection are dissected as parallel tasks, so a BFS may
    // read the mark of a vertex another task is writing (never its own tag)
    std::unique_ptr<std::atomic<int32_t>[]> stamp;
    std::vector<int32_t> level;   // written and read only by the task owning the vertex
    std::atomic<int32_t> cur{0};
    int32_t leaf;
    NdWork(int32_t n, const std::vector<int64_t>& p, const std::vector<int32_t>& a, int32_t lf)
        : ptr(p), adj(a), stamp(new std::atomic<int32_t>[n]), level(n, -1), leaf(lf) {
        for (int32_t i = 0; i < n; ++i) stamp[i].store(-1, std::memory_order_relaxed);
    }
    int32_t ld(int32_t v) const { return stamp[v].load(std::memory_order_relaxed); }
    void st(int32_t v, int32_t t) { stamp[v].store(t, std::memory_order_relaxed); }
    int32_t new_tag() { return cur.fetch_add(2) + 2; }

    // AMD on the subgraph induced by `set`; loc is a scratch map (all -1 on
    // entry and exit) private to the calling thread
    void amd_job(std::vector<int32_t>& set, std::vector<int32_t>& loc) const {
        const int32_t k = (int32_t)set.size();
        if (k <= 1) return;
        for (int32_t i = 0; i < k; ++i) loc[set[i]] = i;
        std::vector<int64_t> lp(k + 1, 0);
        std::vector<int32_t> la;
        for (int32_t i = 0; i < k; ++i) {
            const int32_t v = set[i];
            for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
                const int32_t w = adj[e];
                if (loc[w] >= 0 && w != v) la.push_back(loc[w]);
            }
            lp[i + 1] = (int64_t)la.size();
        }
        std::vector<int32_t> o = amd_order(k, lp, la);
        for (int32_t i = 0; i < k; ++i) loc[set[i]] = -1;
        std::vector<int32_t> out(k);
        for (int32_t i = 0; i < k; ++i) out[i] = set[o[i]];
        set.swap(out);
    }

    // BFS inside the vertices stamped `tag` from `root`; returns the depth
    // (level[] set for the reached vertices) and the visit order
    int32_t bfs(int32_t root, int32_t tag, std::vector<int32_t>& visit) {
        visit.clear();
        visit.push_back(root);
        level[root] = 0;
        st(root, tag + 1);   // reached mark for this BFS
        int32_t depth = 0;
        for (size_t h = 0; h < visit.size(); ++h) {
            const int32_t v = visit[h];
            for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
                const int32_t w = adj[e];
                if (ld(w) == tag) {
                    st(w, tag + 1);
                    level[w] = level[v] + 1;
                    depth = std::max(depth, level[w]);
                    visit.push_back(w);
                }
            }
        }
        for (int32_t v : visit) st(v, tag);   // restore membership
        return depth;
    }

    // the sets to number by minimum degree for the subgraph `set`, in
    // elimination order, appended to `out`
    void dissect(std::vector<int32_t> set, NdJobs& out) {
        if ((int32_t)set.size() <= leaf) { if (!set.empty()) out.push_back(std::move(set)); return; }
        const int32_t tag = new_tag();
        for (int32_t v : set) st(v, tag);
        // the first level structure, from set[0]; when it misses vertices the
        // subgraph is disconnected and its components are dissected on their
        // own (the first component is this BFS's vertex set, in this order)
        std::vector<int32_t> visit;
        int32_t depth = bfs(set[0], tag, visit);
        if (visit.size() < set.size()) {
            NdJobs comps;
            size_t covered = 0;
            const int32_t seen = new_tag();
            for (int32_t v : set) {
                if (ld(v) != tag) continue;
                visit.clear();
                visit.push_back(v);
                st(v, seen);
                for (size_t h = 0; h < visit.size(); ++h) {
                    const int32_t u = visit[h];
                    for (int64_t e = ptr[u]; e < ptr[u + 1]; ++e)
                        if (ld(adj[e]) == tag) { st(adj[e], seen); visit.push_back(adj[e]); }
                }
                covered += visit.size();
                comps.push_back(visit);
                if (covered == set.size()) break;
            }
            for (auto& c : comps) dissect(std::move(c), out);
            return;
        }
        // pseudo-peripheral root: repeat BFS from a minimum-degree vertex of
        // the last level while the depth grows (at most 4 sweeps)
        std::vector<int32_t> lsave;
        for (int it = 0; it < 4; ++it) {
            int32_t best = -1;
            int64_t bdeg = INT64_MAX;
            for (int32_t v : visit)
                if (level[v] == depth) {
                    const int64_t dg = ptr[v + 1] - ptr[v];
                    if (dg < bdeg) { bdeg = dg; best = v; }
                }
            lsave.resize(visit.size());
            for (size_t q = 0; q < visit.size(); ++q) lsave[q] = level[visit[q]];
            std::vector<int32_t> vis2;
            const int32_t d2 = bfs(best, tag, vis2);
            if (d2 < depth) {   // the previous level structure stays
                for (size_t q = 0; q < visit.size(); ++q) level[visit[q]] = lsave[q];
                break;
            }
            const bool grew = d2 > depth;
            depth = d2;
            visit.swap(vis2);
            if (!grew) break;
        }
        if (depth < 2) { out.push_back(std::move(set)); return; }
        std::vector<int64_t> cnt(depth + 1, 0);
        for (int32_t v : visit) cnt[level[v]]++;
        // separator level: the smallest level whose both sides hold at least
        // 30% of the vertices (ties to the one nearest the middle)
        const int64_t tot = (int64_t)visit.size();
        int64_t before = 0;
        int32_t sl = -1;
        int64_t sbest = INT64_MAX, mid_dist = INT64_MAX;
        for (int32_t l = 0; l <= depth; ++l) {
            const int64_t after = tot - before - cnt[l];
            if (l > 0 && l < depth && 10 * before >= 3 * tot && 10 * after >= 3 * tot) {
                const int64_t md = std::llabs(2 * before + cnt[l] - tot);
                if (cnt[l] < sbest || (cnt[l] == sbest && md < mid_dist)) { sbest = cnt[l]; sl = l; mid_dist = md; }
            }
            before += cnt[l];
        }
        if (sl < 0) {   // no balanced level: take the median level
            before = 0;
            for (int32_t l = 0; l <= depth; ++l) {
                if (2 * (before + cnt[l]) >= tot) { sl = std::min(std::max(l, 1), depth - 1); break; }
                before += cnt[l];
            }
        }
        // sides: 0 = near (level < sl), 1 = separator, 2 = far (level > sl);
        // separator vertices without a far neighbour move to the near side,
        // then those without a near neighbour move to the far side
        std::vector<int32_t> sep, near_side, far_side;
        for (int32_t v : visit)
            if (level[v] == sl) sep.push_back(v);
        auto side = [&](int32_t v) { return level[v] < sl ? 0 : (level[v] == sl ? 1 : 2); };
        std::vector<int32_t> keep;
        for (int32_t v : sep) {
            bool far = false;
            for (int64_t e = ptr[v]; e < ptr[v + 1] && !far; ++e)
                far = ld(adj[e]) == tag && side(adj[e]) == 2;
            if (!far) level[v] = sl - 1;   // near side
            else keep.push_back(v);
        }
        sep.clear();
        for (int32_t v : keep) {
            bool nearn = false;
            for (int64_t e = ptr[v]; e < ptr[v + 1] && !nearn; ++e)
                nearn = ld(adj[e]) == tag && side(adj[e]) == 0;
            if (!nearn) level[v] = sl + 1;   // far side
            else sep.push_back(v);
        }
        if (2 * sep.size() > set.size()) { out.push_back(std::move(set)); return; }
        for (int32_t v : visit) {
            const int s = side(v);
            if (s == 0) near_side.push_back(v);
            else if (s == 2) far_side.push_back(v);
        }
        std::sort(near_side.begin(), near_side.end());
        std::sort(far_side.begin(), far_side.end());
        std::sort(sep.begin(), sep.end());
        const bool par = set.size() > 20000;
        set.clear();
        set.shrink_to_fit();
        NdJobs jn, jf;
#pragma omp task shared(jn) if (par)
        dissect(std::move(near_side), jn);
        dissect(std::move(far_side), jf);
#pragma omp taskwait
        for (auto& j : jn) out.push_back(std::move(j));
        for (auto& j : jf) out.push_back(std::move(j));
        out.push_back(std::move(sep));
    }
};

}  // namespace

std::vector<int32_t> nd_order(int32_t n, const std::vector<int64_t>& adjptr,
                              const std::vector<int32_t>& adj, int32_t leaf) {
    NdWork w(n, adjptr, adj, std::max(leaf, 16));
    std::vector<int32_t> all(n);
    for (int32_t i = 0; i < n; ++i) all[i] = i;
    NdJobs jobs;
#pragma omp parallel
#pragma omp single
    w.dissect(std::move(all), jobs);
    const int64_t nj = (int64_t)jobs.size();
#pragma omp parallel
    {
        std::vector<int32_t> loc(n, -1);
#pragma omp for schedule(dynamic, 1)
        for (int64_t j = 0; j < nj; ++j) w.amd_job(jobs[j], loc);
    }
    std::vector<int32_t> order;
    order.reserve(n);
    for (const auto& j : jobs) order.insert(order.end(), j.begin(), j.end());
    return order;
}

std::vector<int32_t> etree(int32_t m, const std::vector<int64_t>& pp, const std::vector<int32_t>& pj) {
    std::vector<int32_t> parent(m, -1), anc(m, -1);
    for (int32_t k = 0; k < m; ++k) {
        for (int64_t t = pp[k]; t < pp[k + 1]; ++t) {
            int32_t i = pj[t];
            while (i != -1 && i < k) {
                int32_t inext = anc[i];
                anc[i] = k;
                if (inext == -1) parent[i] = k;
                i = inext;
            }
        }
    }
    return parent;
}

std::vector<int32_t> postorder(const std::vector<int32_t>& parent) {
    int32_t m = (int32_t)parent.size();
    std::vector<int32_t> head(m, -1), next(m, -1), post;
    post.reserve(m);
    for (int32_t j = m - 1; j >= 0; --j)
        if (parent[j] >= 0) { next[j] = head[parent[j]]; head[parent[j]] = j; }
    std::vector<int32_t> stack;
    for (int32_t r = 0; r < m; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            int32_t v = stack.back();
            int32_t c = head[v];
            if (c == -1) { stack.pop_back(); post.push_back(v); }
            else { head[v] = next[c]; stack.push_back(c); }
        }
    }
    return post;
}

// Column counts of L for a postordered elimination tree (node k of the
// permuted matrix has postorder number k), by the row-subtree leaf method of
// Gilbert, Ng and Peyton (1994) in O(nnz(M) alpha):
//   cc[j] = #{rows i : j in T_i}, T_i = the row subtree of row i, and
//   1[j in T_i] = sum over x in subtree(j) of delta_i(x), with delta_i = +1 at
//   every leaf of T_i, -1 at the least common ancestor of consecutive leaves
//   (postorder), -1 at parent(i), and +1 at i when row i has no entry left
//   of the diagonal.
// A neighbour j < i of row i (taken in increasing j) is a leaf of T_i exactly
// when first[j] (the smallest node of subtree(j)) exceeds the previous
// neighbour of row i.  The least common ancestor of the previous leaf and j
// is found in a disjoint-set forest where every finished node is linked to
// its parent.
std::vector<int32_t> colcounts(int32_t m, const std::vector<int64_t>& pp, const std::vector<int32_t>& pj,
                               const std::vector<int32_t>& parent) {
    for (int32_t j = 0; j < m; ++j)
        if (parent[j] != -1 && parent[j] <= j) throw std::runtime_error("colcounts: tree not postordered");
    std::vector<int32_t> first(m, -1);
    for (int32_t k = 0; k < m; ++k)
        for (int32_t x = k; x != -1 && first[x] == -1; x = parent[x]) first[x] = k;
    std::vector<int32_t> delta(m, 0), prevnbr(m, -1), prevleaf(m, -1), uf(m);
    for (int32_t x = 0; x < m; ++x) uf[x] = x;
    auto find = [&](int32_t x) {
        int32_t r = x;
        while (uf[r] != r) r = uf[r];
        while (uf[x] != r) { int32_t nx = uf[x]; uf[x] = r; x = nx; }
        return r;
    };
    for (int32_t j = 0; j < m; ++j) {
        for (int64_t t = pp[j]; t < pp[j + 1]; ++t) {
            const int32_t i = pj[t];
            if (i <= j || prevnbr[i] >= first[j]) { if (i > j) prevnbr[i] = j; continue; }
            delta[j] += 1;
            if (prevleaf[i] != -1) delta[find(prevleaf[i])] -= 1;
            prevleaf[i] = j;
            prevnbr[i] = j;
        }
        if (parent[j] != -1) uf[j] = parent[j];
    }
    for (int32_t i = 0; i < m; ++i) {
        if (parent[i] != -1) delta[parent[i]] -= 1;
        if (prevleaf[i] == -1) delta[i] += 1;
    }
    for (int32_t j = 0; j < m; ++j)
        if (parent[j] != -1) delta[parent[j]] += delta[j];
    return delta;
}

// Reference form (row-subtree walk, O(nnz(L))), kept for the tests.
std::vector<int32_t> colcounts_walk(int32_t m, const std::vector<int64_t>& adjptr,
                                    const std::vector<int32_t>& adj, const std::vector<int32_t>& perm,
                                    const std::vector<int32_t>& iperm, const std::vector<int32_t>& parent) {
    std::vector<int32_t> cc(m, 1), mark(m, -1);
    for (int32_t i = 0; i < m; ++i) {   // row subtree of row i of L
        mark[i] = i;
        int32_t old = perm[i];
        for (int64_t t = adjptr[old]; t < adjptr[old + 1]; ++t) {
            int32_t k = iperm[adj[t]];
            if (k >= i) continue;
            while (mark[k] != i) { cc[k]++; mark[k] = i; k = parent[k]; }
        }
    }
    return cc;
}

Supernodal analyze(int32_t m, const std::vector<int64_t>& adjptr, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& perm_in, const RelaxParams& rp) {
    Supernodal S;
    S.m = m;
    std::vector<int32_t> perm = perm_in.empty() ? amd_order(m, adjptr, adj) : perm_in;
    std::vector<int32_t> iperm(m);
    for (int32_t k = 0; k < m; ++k) iperm[perm[k]] = k;
    // etree, then postorder it: a postorder is a topological order of the
    // tree, so the permuted matrix has the same fill and the same tree with
    // its nodes relabelled
    // (the graph relabelled once per numbering: the sequential sweeps below
    // then read it in order)
    std::vector<int64_t> pp;
    std::vector<int32_t> pj;
    permute_graph(m, adjptr, adj, perm, iperm, pp, pj, -1);
    std::vector<int32_t> par = etree(m, pp, pj);
    std::vector<int32_t> post = postorder(par);
    std::vector<int32_t> perm2(m), ipost(m);
    for (int32_t k = 0; k < m; ++k) { perm2[k] = perm[post[k]]; ipost[post[k]] = k; }
    perm.swap(perm2);
    for (int32_t k = 0; k < m; ++k) iperm[perm[k]] = k;
    std::vector<int32_t> par2(m);
    for (int32_t k = 0; k < m; ++k) par2[k] = par[post[k]] >= 0 ? ipost[par[post[k]]] : -1;
    par.swap(par2);
    permute_graph(m, adjptr, adj, perm, iperm, pp, pj, 1);
    std::vector<int32_t> cc = colcounts(m, pp, pj, par);
    S.perm = perm; S.iperm = iperm; S.parent = par; S.colcount = cc;
    for (int32_t j = 0; j < m; ++j) {
        S.nnz_l_exact += cc[j];
        S.flops += (double)cc[j] * (double)cc[j];
    }
    // fundamental supernodes
    std::vector<int32_t> nchild(m, 0);
    for (int32_t j = 0; j < m; ++j) if (par[j] >= 0) nchild[par[j]]++;
    std::vector<int32_t> fs{0};
    for (int32_t j = 1; j < m; ++j)
        if (!(par[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1)) fs.push_back(j);
    fs.push_back(m);
    int32_t nf = (int32_t)fs.size() - 1;
    // relaxed amalgamation, bottom-up: a group may absorb into its parent
    // group when the parent starts right after it (it is the last child).
    std::vector<int32_t> fcol2sn(m);
    for (int32_t s = 0; s < nf; ++s)
        for (int32_t j = fs[s]; j < fs[s + 1]; ++j) fcol2sn[j] = s;
    std::vector<int32_t> gfirst(nf), gncol(nf), gnrow(nf);
    std::vector<int64_t> gactual(nf);
    std::vector<int32_t> merged_into(nf, -1);
    for (int32_t s = 0; s < nf; ++s) {
        gfirst[s] = fs[s];
        gncol[s] = fs[s + 1] - fs[s];
        gnrow[s] = cc[fs[s]];
        int64_t a = 0;
        for (int32_t j = fs[s]; j < fs[s + 1]; ++j) a += cc[j];
        gactual[s] = a;
    }
    for (int32_t s = 0; s < nf; ++s) {
        int32_t last = fs[s + 1] - 1;
        int32_t pc = par[last];
        if (pc < 0) continue;
        int32_t p = fcol2sn[pc];
        if (p != s + 1) continue;                 // not contiguous: cannot merge
        int64_t nc = (int64_t)gncol[s] + gncol[p];
        if (nc > rp.max_cols) continue;
        int64_t nr = (int64_t)gncol[s] + gnrow[p];
        int64_t stored = nc * nr - nc * (nc - 1) / 2;
        int64_t zeros = stored - (gactual[s] + gactual[p]);
        double frac = stored > 0 ? (double)zeros / (double)stored : 0.0;
        bool ok = nc <= rp.nrelax[0] ||
                  (nc <= rp.nrelax[1] && frac < rp.zrelax[0]) ||
                  (nc <= rp.nrelax[2] && frac < rp.zrelax[1]) ||
                  frac < rp.zrelax[2];
        if (!ok) continue;
        gfirst[p] = gfirst[s];
        gncol[p] = (int32_t)nc;
        gnrow[p] = (int32_t)nr;
        gactual[p] += gactual[s];
        merged_into[s] = p;
    }
    S.sfirst.clear();
    for (int32_t s = 0; s < nf; ++s) {
        if (merged_into[s] >= 0) continue;
        // a supernode split into consecutive pieces stays a valid partition:
        // each piece's parent is the next piece
        int32_t f = gfirst[s], l = fs[s + 1];
        for (int32_t k = f; k < l; k += rp.max_sn_cols) S.sfirst.push_back(k);
    }
    S.sfirst.push_back(m);
    S.nsuper = (int32_t)S.sfirst.size() - 1;
    const int32_t ns = S.nsuper;
    S.col2sn.resize(m);
    for (int32_t s = 0; s < ns; ++s)
        for (int32_t j = S.sfirst[s]; j < S.sfirst[s + 1]; ++j) S.col2sn[j] = s;
    S.sparent.assign(ns, -1);
    for (int32_t s = 0; s < ns; ++s) {
        int32_t pc = par[S.sfirst[s + 1] - 1];
        S.sparent[s] = pc < 0 ? -1 : S.col2sn[pc];
    }
    S.childptr.assign(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) if (S.sparent[s] >= 0) S.childptr[S.sparent[s] + 1]++;
    for (int32_t s = 0; s < ns; ++s) S.childptr[s + 1] += S.childptr[s];
    S.child.resize(S.childptr[ns]);
    {
        std::vector<int32_t> f(S.childptr.begin(), S.childptr.end() - 1);
        for (int32_t s = 0; s < ns; ++s) if (S.sparent[s] >= 0) S.child[f[S.sparent[s]]++] = s;
    }
    // row structures: own columns U lower entries of M U children's update rows
    S.rowoff.assign(ns + 1, 0);
    S.snrow.assign(ns, 0);
    std::vector<std::vector<int32_t>> tmp(ns);
    {
        // a supernode's structure needs its children's: one parallel sweep
        // per elimination-tree level (children sit on lower levels)
        std::vector<int32_t> lev(ns, 0);
        int32_t top = 0;
        for (int32_t s = 0; s < ns; ++s) {
            for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) lev[s] = std::max(lev[s], lev[S.child[t]] + 1);
            top = std::max(top, lev[s]);
        }
        std::vector<int32_t> lptr(top + 2, 0), lsn(ns);
        for (int32_t s = 0; s < ns; ++s) lptr[lev[s] + 1]++;
        for (int32_t l = 0; l <= top; ++l) lptr[l + 1] += lptr[l];
        {
            std::vector<int32_t> f(lptr.begin(), lptr.end() - 1);
            for (int32_t s = 0; s < ns; ++s) lsn[f[lev[s]]++] = s;
        }
#pragma omp parallel
        {
            std::vector<int32_t> mk(m, -1), rows;
            for (int32_t l = 0; l <= top; ++l) {
#pragma omp for schedule(dynamic, 16)
                for (int32_t q = lptr[l]; q < lptr[l + 1]; ++q) {
                    const int32_t s = lsn[q];
                    const int32_t f = S.sfirst[s], lc = S.sfirst[s + 1];
                    rows.clear();
                    for (int32_t j = f; j < lc; ++j) { mk[j] = s; rows.push_back(j); }
                    for (int32_t j = f; j < lc; ++j)
                        for (int64_t t = pp[j]; t < pp[j + 1]; ++t) {
                            const int32_t i = pj[t];
                            if (i >= lc && mk[i] != s) { mk[i] = s; rows.push_back(i); }
                        }
                    for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) {
                        const int32_t c = S.child[t];
                        const int32_t cn = S.sfirst[c + 1] - S.sfirst[c];
                        const std::vector<int32_t>& cr = tmp[c];
                        for (size_t u = cn; u < cr.size(); ++u)
                            if (mk[cr[u]] != s) { mk[cr[u]] = s; rows.push_back(cr[u]); }
                    }
                    std::sort(rows.begin(), rows.end());
                    tmp[s] = rows;
                }
            }
        }
    }
    for (int32_t s = 0; s < ns; ++s) {
        S.snrow[s] = (int32_t)tmp[s].size();
        S.rowoff[s + 1] = S.rowoff[s] + (int64_t)tmp[s].size();
    }
    S.rowidx.resize(S.rowoff[ns]);
    for (int32_t s = 0; s < ns; ++s) {
        std::copy(tmp[s].begin(), tmp[s].end(), S.rowidx.begin() + S.rowoff[s]);
        std::vector<int32_t>().swap(tmp[s]);
    }
    S.loff.assign(ns + 1, 0);
    S.uoff.assign(ns + 1, 0);
    S.reloff.assign(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) {
        int64_t nc = S.sfirst[s + 1] - S.sfirst[s], nr = S.snrow[s];
        S.loff[s + 1] = S.loff[s] + nc * nr;
        S.uoff[s + 1] = S.uoff[s] + (nr - nc) * (nr - nc);
        S.reloff[s + 1] = S.reloff[s] + (nr - nc);
    }
    S.relmap.assign(S.reloff[ns], -1);
    std::vector<int32_t> pos(m, -1);
    for (int32_t p = 0; p < ns; ++p) {
        for (int64_t t = S.rowoff[p]; t < S.rowoff[p + 1]; ++t) pos[S.rowidx[t]] = (int32_t)(t - S.rowoff[p]);
        for (int32_t t = S.childptr[p]; t < S.childptr[p + 1]; ++t) {
            int32_t c = S.child[t];
            int32_t cn = S.sfirst[c + 1] - S.sfirst[c];
            for (int32_t u = cn; u < S.snrow[c]; ++u) {
                int32_t r = S.rowidx[S.rowoff[c] + u];
                S.relmap[S.reloff[c] + u - cn] = pos[r];
            }
        }
    }
    // level schedule: leaves at level 0, parent after all children
    S.level.assign(ns, 0);
    int32_t maxl = 0;
    for (int32_t s = 0; s < ns; ++s) {
        int32_t lv = 0;
        for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) lv = std::max(lv, S.level[S.child[t]] + 1);
        S.level[s] = lv;
        maxl = std::max(maxl, lv);
    }
    S.nlevels = ns ? maxl + 1 : 0;
    S.levelptr.assign(S.nlevels + 1, 0);
    for (int32_t s = 0; s < ns; ++s) S.levelptr[S.level[s] + 1]++;
    for (int32_t l = 0; l < S.nlevels; ++l) S.levelptr[l + 1] += S.levelptr[l];
    S.levelsn.resize(ns);
    {
        std::vector<int32_t> f(S.levelptr.begin(), S.levelptr.end() - 1);
        for (int32_t s = 0; s < ns; ++s) S.levelsn[f[S.level[s]]++] = s;
    }
    return S;
}

// Where each ADA^T slot lands in the supernodal factor (slot_dest), the
// diagonal of every row (diagpos; also for an empty row of A, whose ADA^T
// row has no slot but still receives beta), the column of every nonzero of
// A (colof) and the values of A in CSR order (csrval).
void assembly_maps(const Supernodal& S, const AdatPattern& P, const CscMatrix& A, const CsrMatrix& R,
                   std::vector<int64_t>& slot_dest, std::vector<int64_t>& diagpos,
                   std::vector<int32_t>& colof, std::vector<double>& csrval) {
    const int32_t m = A.m, n = A.n;
    const int64_t nnz = (int64_t)A.rowind.size();
    slot_dest.resize(P.colind.size());
    diagpos.resize(m);
    int64_t missing = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : missing)
    for (int32_t i = 0; i < m; ++i)
        for (int64_t s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
            const int32_t k = P.colind[s];
            const int32_t a = S.iperm[i], b = S.iperm[k];
            const int32_t col = std::min(a, b), row = std::max(a, b);
            const int32_t sn = S.col2sn[col];
            const int32_t* rb = S.rowidx.data() + S.rowoff[sn];
            const int32_t* re = S.rowidx.data() + S.rowoff[sn + 1];
            const int32_t* it = std::lower_bound(rb, re, row);
            if (it == re || *it != row) { ++missing; slot_dest[s] = 0; continue; }
            slot_dest[s] = S.loff[sn] + (int64_t)(col - S.sfirst[sn]) * S.snrow[sn] + (it - rb);
        }
    if (missing) throw std::runtime_error("assembly map: row not in structure");
    // own columns lead a supernode's row structure
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < m; ++i) {
        const int32_t col = S.iperm[i], sn = S.col2sn[col];
        const int64_t r = col - S.sfirst[sn];
        diagpos[i] = S.loff[sn] + r * S.snrow[sn] + r;
    }
    colof.resize(nnz);
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < n; ++j)
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) colof[p] = j;
    csrval.resize(nnz);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nnz; ++t) csrval[t] = A.val[R.cscpos[t]];
}

}  // namespace help
