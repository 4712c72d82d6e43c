// MI355X normal-equation backend: handle, schedules, renew/solve, C ABI.
//
// help_gpu_create   <- LinearEquation_Construct (UseCholmod.cpp:29-58)
// help_gpu_renew    <- RenewLinearEquation      (UseCholmod.cpp:84-149)
// help_gpu_solve    <- SolveLinearEquation      (UseCholmod.cpp:157-238)
// help_gpu_destroy  <- LinearEquation_Destruct  (UseCholmod.cpp:60-81)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <exception>
#include <thread>
#include <type_traits>
#include <vector>

#include "device.h"
#include "help_mi355x.h"
#include "symbolic.h"

using namespace help;

#include "handle.h"

namespace help {
thread_local std::string g_create_error;
}
using help::g_create_error;
using help::now_s;
using help::kKindName;

namespace {

// Position of slot (i,k) (original indices) inside the supernodal L storage.
void build_assembly_maps(help_gpu* h) {
    help::assembly_maps(h->S, h->P, h->A, h->R, h->slot_dest, h->diagpos, h->colof, h->csrval);
    h->csrcol = h->R.colind;
}

// ---- profiling: events around launches on the handle's stream ----
struct ProfScope {
    help_gpu* h;
    int32_t kind;
    double work;
    hipEvent_t b = nullptr;
    size_t idx = 0;
    ProfScope(help_gpu* h_, int32_t k, double w) : h(h_), kind(k), work(w) {
        if (!h->prof) return;
        size_t used = h->pending.size() * 2;
        while (h->evpool.size() < used + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            h->evpool.push_back(e);
        }
        hipEvent_t a = h->evpool[used];
        b = h->evpool[used + 1];
        idx = h->pending.size();
        h->pending.push_back({kind, work, a, b, h->prof_fp32});
        // the launches of the region take a and b as launch events
        help::g_region_events = {a, b, false};
    }
    // work known only after the launches (iterative kernels)
    void set_work(double w) {
        if (h->prof && b) h->pending[idx].work = w;
    }
    ~ProfScope() {
        if (!h->prof || !b) return;
        if (!help::g_region_events.used) {
            // nothing was launched: an empty interval
            (void)hipEventRecord(h->pending[idx].a, h->st);
            (void)hipEventRecord(b, h->st);
        }
        help::g_region_events = {};
    }
};

// after the stream has been synchronized
void collect_prof(help_gpu* h) {
    for (auto& p : h->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            for (help_gpu::KStat* k : {&h->kstat[p.kind], p.fp32 ? &h->kstat32[p.kind] : nullptr}) {
                if (!k) continue;
                k->ms += ms;
                k->work += p.work;
                k->launches += 1;
            }
        }
    }
    h->pending.clear();
}

// lower-triangle entries (i >= j) of a 64 x 64 tile at (r0, c0) below
// row/col limits: sum over its columns j of the rows max(r0, j) .. hi-1
double tile_lower_entries(int32_t r0, int32_t c0, int32_t rlim, int32_t clim) {
    const int64_t hi = std::min(r0 + kTile, rlim), jend = std::min(c0 + kTile, clim);
    double cnt = 0;
    // columns left of r0 see the whole row range
    const int64_t jl = std::min<int64_t>(jend, r0);
    if (jl > c0 && hi > r0) cnt += (double)(jl - c0) * (double)(hi - r0);
    // columns j >= r0 see hi - j rows while j < hi
    const int64_t ja = std::max<int64_t>(c0, r0), jb = std::min<int64_t>(jend, hi);
    if (jb > ja) cnt += (double)(jb - ja) * (double)(2 * hi - ja - jb + 1) / 2.0;
    return cnt;
}

bool is_small_front(int32_t c, int32_t ncol) {
    return c <= kSmallRows && (int64_t)c * ncol <= kSmallElems;
}

bool is_big_solve(int32_t c, int32_t ncol) {
    return (c - ncol) > 128 && (int64_t)ncol * (c - ncol) > 32768;
}

void build_schedules(help_gpu* h) {
    const Supernodal& S = h->S;
    h->sbig.assign(S.nsuper, 0);
    auto ncol = [&](int32_t s) { return S.sfirst[s + 1] - S.sfirst[s]; };
    // explicit inverses of L11 for supernodes wider than one 64-column block
    h->use_xinv = h->opts.inverse_solve >= 0;
    auto nblk_of = [&](int32_t s) { return (ncol(s) + kTile - 1) / kTile; };
    auto wide = [&](int32_t s) { return h->use_xinv && nblk_of(s) > 1; };
    h->xoff.assign(S.nsuper + 1, 0);
    for (int32_t s = 0; s < S.nsuper; ++s) {
        const int64_t nb = nblk_of(s);
        h->xoff[s + 1] = h->xoff[s] + (wide(s) ? nb * (nb - 1) / 2 * kTile * kTile : 0);
    }
    // diagonal-block inverses: 64x64 blocks, the last one nb x nb
    h->dinvoff.assign(S.nsuper + 1, 0);
    for (int32_t s = 0; s < S.nsuper; ++s) {
        int64_t nc = ncol(s), full = nc / kTile, rem = nc % kTile;
        h->dinvoff[s + 1] = h->dinvoff[s] + full * kTile * kTile + rem * rem;
    }
    // Cholesky flops of a supernode's columns: sum over its columns of colcount^2
    auto sn_flops = [&](int32_t s) {
        double f = 0;
        for (int32_t k = 0; k < ncol(s); ++k) { double cj = S.snrow[s] - k; f += cj * cj; }
        return f;
    };
    // bytes of reading the children's update matrices and updating the front
    auto ea_bytes = [&](int32_t s) {
        double b = 0;
        for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) {
            double cu = S.snrow[S.child[t]] - ncol(S.child[t]);
            b += 3.0 * 8.0 * cu * (cu + 1) / 2;
        }
        return b;
    };
    // Phases, as soon as possible.  A small front is one SMALL launch; a
    // large front with P panels takes phases start .. start+P-1: extend-add
    // and panel 0 in the first, panel b in phase start+b, and the update
    // matrix U -= L21 L21^T in the last.  A front starts once its children's
    // update matrices are complete: the phase after a large child's last
    // one; a small child's own phase when the parent is large (its SMALL
    // launch comes first within the phase), the next one when both are small.
    // Supernodes are postordered, so children come before their parent.
    // Within a phase the launches run SMALL, EA, POTRF, TRSM, GEMM (in-panel
    // trailing updates), GEMM (update matrices), then the side-stream XINV
    // launches of the wide fronts (see below).
    const int32_t ns = S.nsuper;
    // A front whose update matrix is small takes its U -= L21 L21^T in the
    // GEMM launch of every panel (the accumulator parked in Uacc between
    // panels) instead of a launch of its own after the last panel.
    auto u_tiles = [&](int32_t s) {
        const int64_t t = (S.snrow[s] - ncol(s) + kTile - 1) / kTile;
        return t * (t + 1) / 2;
    };
    auto folded = [&](int32_t s, int32_t panels) {
        const int64_t ut = u_tiles(s);
        return ut > 0 && ut <= 256 && ut * panels <= 1024;
    };
    h->use_uacc = false;
    std::vector<int32_t> start(ns), done(ns), npan(ns);
    std::vector<char> smallf(ns);
    int32_t nphase = 0;
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t nc = ncol(s), c = S.snrow[s];
        smallf[s] = h->opts.front_path != 1 && is_small_front(c, nc);
        npan[s] = smallf[s] ? 1 : (nc + kTile - 1) / kTile;
        int32_t r = 0;
        for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) {
            const int32_t ch = S.child[t];
            r = std::max(r, done[ch] + ((smallf[ch] && !smallf[s]) ? 0 : 1));
        }
        start[s] = r;
        done[s] = r + npan[s] - 1;
        nphase = std::max(nphase, done[s] + 1);
    }
    // Update-matrix placement.  Front s's update matrix is live from its
    // first phase (its extend-add, or its small-front launch, zeroes it)
    // through its parent's first phase (the parent's extend-add reads it),
    // both inclusive; fronts whose lifetimes do not overlap share memory.
    // First fit over an address-ordered free list, phase by phase: the blocks
    // that died before the phase are freed, then its new ones are placed.
    {
        h->uplan.assign(ns + 1, 0);
        std::vector<std::vector<int32_t>> born(nphase + 1), dies(nphase + 1);
        for (int32_t s = 0; s < ns; ++s) {
            const int64_t cu = S.snrow[s] - ncol(s);
            if (cu <= 0) continue;
            born[start[s]].push_back(s);
            const int32_t last = S.sparent[s] >= 0 ? start[S.sparent[s]] : done[s];
            dies[last + 1].push_back(s);
        }
        auto usize = [&](int32_t s) {
            const int64_t cu = S.snrow[s] - ncol(s);
            return (cu * cu + 63) / 64 * 64;   // 64-element aligned blocks
        };
        std::map<int64_t, int64_t> freel;   // offset -> length
        int64_t top = 0, high = 0;
        for (int32_t p = 0; p <= nphase; ++p) {
            for (int32_t s : dies[p]) {
                int64_t o = h->uplan[s], len = usize(s);
                auto nx = freel.lower_bound(o);
                if (nx != freel.end() && nx->first == o + len) { len += nx->second; nx = freel.erase(nx); }
                if (nx != freel.begin()) {
                    auto pv = std::prev(nx);
                    if (pv->first + pv->second == o) { o = pv->first; len += pv->second; freel.erase(pv); }
                }
                if (o + len == top) top = o;
                else freel[o] = len;
            }
            if (p == nphase) break;
            for (int32_t s : born[p]) {
                const int64_t len = usize(s);
                auto it = freel.begin();
                while (it != freel.end() && it->second < len) ++it;
                if (it != freel.end()) {
                    h->uplan[s] = it->first;
                    if (it->second > len) freel[it->first + len] = it->second - len;
                    freel.erase(it);
                } else {
                    // the free block just below the top, if any, grows into it
                    int64_t o = top;
                    if (!freel.empty()) {
                        auto last = std::prev(freel.end());
                        if (last->first + last->second == top) { o = last->first; freel.erase(last); }
                    }
                    h->uplan[s] = o;
                    top = o + len;
                    high = std::max(high, top);
                }
            }
        }
        h->uplan[ns] = high;
        if (h->opts.verbose)
            std::fprintf(stderr, "[help] update matrices: %lld elements placed (%lld without reuse)\n",
                         (long long)high, (long long)S.uoff[ns]);
    }
    // fronts by phase: small ones, large ones starting, large ones active
    // (some panel in the phase), large ones finishing
    std::vector<std::vector<int32_t>> at_small(nphase), at_start(nphase), at_active(nphase), at_done(nphase);
    for (int32_t s = 0; s < ns; ++s) {
        if (smallf[s]) { at_small[start[s]].push_back(s); continue; }
        at_start[start[s]].push_back(s);
        for (int32_t p = start[s]; p <= done[s]; ++p) at_active[p].push_back(s);
        at_done[done[s]].push_back(s);
    }
    for (int32_t p = 0; p < nphase; ++p) {
        if (!at_small[p].empty()) {
            // fronts of at most 64 rows and 16 columns whose panel and
            // inverse scratch fit kSmallWaveLds take one wave each, the others
            // a workgroup; only in phases with enough such fronts to fill the
            // chip (a lone one-wave front walks its columns serially per lane
            // and is slower than the workgroup form)
            auto wave_ok = [&](int32_t s) {
                const int32_t nc = ncol(s), c = S.snrow[s];
                const int32_t nbm = std::min(kTile, nc);
                return c <= 64 && nc <= 16 && c * nc + nbm * (nbm + 1) <= kSmallWaveLds;
            };
            int32_t nwave = 0;
            for (int32_t s : at_small[p]) nwave += wave_ok(s) ? 1 : 0;
            const bool force64 = h->opts.front_path == 2;
            const bool use_wave = nwave >= 2048 || (force64 && nwave > 0);
            for (int pass = 0; pass < 2; ++pass) {
                int32_t off = (int32_t)h->list_small.size();
                double w = 0;
                int32_t lds = 0;
                for (int32_t s : at_small[p]) {
                    const int32_t nc = ncol(s), c = S.snrow[s];
                    const int32_t nbm = std::min(kTile, nc);
                    const int32_t need = c * nc + nbm * (nbm + 1);
                    const bool one_wave = use_wave && wave_ok(s);
                    if (one_wave != (pass == 0)) continue;
                    h->list_small.push_back(s);
                    w += sn_flops(s);
                    lds = std::max(lds, need);
                }
                if ((int32_t)h->list_small.size() > off) {
                    Op o{OP_SMALL, off, (int32_t)h->list_small.size() - off, w, lds};
                    if (pass == 0) {   // 32 lanes per front when every front has at most 32 rows
                        o.wave = force64 ? 64 : 32;
                        for (int32_t t = off; t < (int32_t)h->list_small.size(); ++t)
                            if (S.snrow[h->list_small[t]] > 32) o.wave = 64;
                    }
                    h->fops.push_back(o);
                }
            }
        }
        if (at_active[p].empty()) continue;
        // extend-add by kEaCols-column chunks of each starting front; the
        // same tasks first zero the front's update matrix (a childless front
        // gets the tasks of its update-matrix columns for that alone)
        int32_t off = (int32_t)h->ea.size();
        double w = 0;
        std::vector<int32_t> chunk_cnt, chunk_fill;
        for (int32_t s : at_start[p]) {
            const int32_t c = S.snrow[s], nc = ncol(s);
            const bool leaf = S.childptr[s + 1] == S.childptr[s];
            if (leaf && c == nc) continue;
            w += ea_bytes(s) + 4.0 * (c - nc) * (c - nc + 1);
            // each child's relative map is ascending: one walk over it cuts it
            // into the kEaCols-column chunks it touches, and the chunks' child
            // lists are filled child by child (child order within a chunk)
            const int32_t nchunk = (c + kEaCols - 1) / kEaCols;
            chunk_cnt.assign(nchunk + 1, 0);
            for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) {
                const int32_t ch = S.child[t];
                const int32_t chc = S.snrow[ch] - ncol(ch);
                const int32_t* rel = S.relmap.data() + S.reloff[ch];
                for (int32_t j = 0; j < chc;) {
                    const int32_t k = rel[j] / kEaCols;
                    while (j < chc && rel[j] / kEaCols == k) ++j;
                    chunk_cnt[k + 1]++;
                }
            }
            for (int32_t k = 0; k < nchunk; ++k) chunk_cnt[k + 1] += chunk_cnt[k];
            const int32_t base = (int32_t)h->earef.size();
            h->earef.resize(base + chunk_cnt[nchunk]);
            chunk_fill.assign(chunk_cnt.begin(), chunk_cnt.end() - 1);
            for (int32_t t = S.childptr[s]; t < S.childptr[s + 1]; ++t) {
                const int32_t ch = S.child[t];
                const int32_t chc = S.snrow[ch] - ncol(ch);
                const int32_t* rel = S.relmap.data() + S.reloff[ch];
                for (int32_t j = 0; j < chc;) {
                    const int32_t k = rel[j] / kEaCols, lo = j;
                    while (j < chc && rel[j] / kEaCols == k) ++j;
                    h->earef[base + chunk_fill[k]++] = {h->uplan[ch], S.reloff[ch], chc, lo, j};
                }
            }
            // a task per chunk with children, and per update-matrix chunk
            // (its columns are zeroed there)
            for (int32_t k = leaf ? nc / kEaCols : 0; k < nchunk; ++k) {
                const int32_t c0 = k * kEaCols, c1 = std::min(c, c0 + kEaCols);
                const int32_t nref = chunk_cnt[k + 1] - chunk_cnt[k];
                if (nref == 0 && c1 <= nc) continue;
                h->ea.push_back({s, c0, c1, base + chunk_cnt[k], nref, nc, c, S.loff[s], h->uplan[s]});
            }
        }
        if ((int32_t)h->ea.size() > off) {
            Op oe{OP_EA, off, (int32_t)h->ea.size() - off, w};
            for (int32_t s : at_start[p]) if (S.snrow[s] >= kEaDeepRows) oe.depth = 8;
            if (h->opts.ea_depth == 4 || h->opts.ea_depth == 8) oe.depth = h->opts.ea_depth;
            h->fops.push_back(oe);
        }
        // the panel of every active front
        int32_t po = (int32_t)h->potrf.size(), to = (int32_t)h->trsm.size(), go = (int32_t)h->gemm.size();
        const int32_t pa = (int32_t)h->panel.size();
        double wp = 0, wt = 0, wg = 0;
        for (int32_t s : at_active[p]) {
            const int32_t nc = ncol(s), c = S.snrow[s];
            const int32_t b = (p - start[s]) * kTile;
            const int32_t nb = std::min(kTile, nc - b);
            h->potrf.push_back({s, b, 0});
            wp += (double)nb * nb * nb / 3.0;
            if (b + nb >= c) h->panel.push_back({s, b, 0});
            for (int32_t r0 = b + nb, rb = 1; r0 < c; r0 += kTile, ++rb) {
                h->trsm.push_back({s, b, rb});
                h->panel.push_back({s, b, rb});
                wt += (double)nb * nb * std::min(kTile, c - r0);
            }
            // Trailing updates by pairs of panels: the first panel of a pair
            // updates only the second one's columns; the second updates every
            // later column with both panels at once (an inner dimension of
            // up to 128), so the trailing tiles are read and written once
            // per two panels.  A front's last panel closes the pair it is in.
            const int32_t pb = p - start[s], P = npan[s];
            const bool pair_first = (pb % 2 == 0) && pb + 1 < P;
            const int32_t kp = (pb % 2 == 1) ? b - kTile : b;   // first panel not yet applied
            for (int32_t c0 = b + nb; c0 < nc; c0 += kTile) {
                if (pair_first && c0 > b + nb) break;
                const int32_t k0 = pair_first ? b : kp;
                for (int32_t r0 = c0; r0 < c; r0 += kTile) {
                    h->gemm.push_back({s, r0, c0, k0, b + nb, nc});
                    wg += 2.0 * (b + nb - k0) * tile_lower_entries(r0, c0, c, nc);
                }
            }
            if (folded(s, npan[s]) && !pair_first) {
                // update points: the second panel of each pair and the last
                const int32_t q = pb / 2, Q = (P + 1) / 2;
                int32_t mode = kGemmNext, k0 = kp, k1 = b + nb;
                if (Q == 1) { mode = kGemmDirect; k0 = 0; k1 = nc; }
                else if (q == 0) mode = kGemmFirst;
                else if (q == Q - 1) mode = kGemmLast;
                if (Q > 1) h->use_uacc = true;
                for (int32_t c0 = nc; c0 < c; c0 += kTile)
                    for (int32_t r0 = c0; r0 < c; r0 += kTile) {
                        GemmTask g{s, r0, c0, k0, k1, c};
                        g.mode = mode;
                        h->gemm.push_back(g);
                        wg += 2.0 * (k1 - k0) * tile_lower_entries(r0, c0, c, c);
                    }
            }
        }
        // FP64: one fused launch for the phase's panels when its workgroups
        // (one per row block, 104 KB of LDS each) fit the chip at once
        const int32_t npa = (int32_t)h->panel.size() - pa;
        const bool fuse = npa <= kFusedPanelMax;
        Op op{OP_POTRF, po, (int32_t)h->potrf.size() - po, wp};
        if (fuse) {
            op.xoff = pa; op.xcount = npa; op.work2 = wp + wt;
            for (int32_t s : at_active[p]) h->pcopy.push_back({s, (p - start[s]) * kTile, 0});
        }
        h->fops.push_back(op);
        if ((int32_t)h->trsm.size() > to) {
            Op ot{OP_TRSM, to, (int32_t)h->trsm.size() - to, wt};
            ot.fused = fuse;
            h->fops.push_back(ot);
        }
        if ((int32_t)h->gemm.size() > go) h->fops.push_back({OP_GEMM, go, (int32_t)h->gemm.size() - go, wg});
        // U -= L21 L21^T with the whole panel as inner dimension, for the
        // fronts whose last panel is this phase's
        go = (int32_t)h->gemm.size();
        wg = 0;
        for (int32_t s : at_done[p]) {
            const int32_t nc = ncol(s), c = S.snrow[s];
            if (folded(s, npan[s])) continue;
            for (int32_t c0 = nc; c0 < c; c0 += kTile)
                for (int32_t r0 = c0; r0 < c; r0 += kTile) {
                    h->gemm.push_back({s, r0, c0, 0, nc, c});
                    wg += 2.0 * nc * tile_lower_entries(r0, c0, c, c);
                }
        }
        if ((int32_t)h->gemm.size() > go) h->fops.push_back({OP_GEMM, go, (int32_t)h->gemm.size() - go, wg});
        // Inverse blocks of the wide fronts, right-looking: task (i, j, k)
        // finalises X_kj (i = k) or adds L_ik X_kj to the partial sum of X_ij
        // (i > k), and every partial sum takes its terms in increasing k.
        // Fronts near the root, whose chain of nb launches after the last
        // panel (about nb/3 phases' time) would outlast the phases left, get
        // the tasks of panel k = p - start[s] as soon as that panel is
        // factored; the others get all their tasks after their last panel,
        // one launch per k - j: fewer launches and fewer forks.
        auto early = [&](int32_t s) { return done[s] + nblk_of(s) / 3 >= nphase - 1; };
        auto task_work = [](int32_t i, int32_t j, int32_t k) {
            return 2.0 * kTile * kTile * kTile * ((k > j ? 1 : 0) + (i > k ? 1 : 0));
        };
        {
            const int32_t xo = (int32_t)h->xtask.size();
            double wx = 0;
            for (int32_t s : at_active[p]) {
                if (!wide(s) || !early(s)) continue;
                const int32_t nb = nblk_of(s), k = p - start[s];
                for (int32_t j = 0; j <= k; ++j)
                    for (int32_t i = k; i < nb; ++i) {
                        if (i == j) continue;   // X_kk is Dinv_k
                        h->xtask.push_back({s, i, j, k});
                        wx += task_work(i, j, k);
                    }
            }
            if ((int32_t)h->xtask.size() > xo)
                h->fops.push_back({OP_XINV, xo, (int32_t)h->xtask.size() - xo, wx});
        }
        std::vector<int32_t> wide_done;
        for (int32_t s : at_done[p])
            if (wide(s) && !early(s)) wide_done.push_back(s);
        int32_t maxnb = 0;
        for (int32_t s : wide_done) maxnb = std::max(maxnb, nblk_of(s));
        for (int32_t step = 0; step < maxnb; ++step) {
            const int32_t xo = (int32_t)h->xtask.size();
            double wx = 0;
            for (int32_t s : wide_done) {
                const int32_t nb = nblk_of(s);
                for (int32_t j = 0; j + step < nb; ++j) {
                    const int32_t k = j + step;
                    for (int32_t i = k; i < nb; ++i) {
                        if (i == j) continue;
                        h->xtask.push_back({s, i, j, k});
                        wx += task_work(i, j, k);
                    }
                }
            }
            if ((int32_t)h->xtask.size() > xo)
                h->fops.push_back({OP_XINV, xo, (int32_t)h->xtask.size() - xo, wx});
        }
    }
    for (GemmTask& g : h->gemm) {
        g.n = ncol(g.s);
        g.c = S.snrow[g.s];
        g.loff = S.loff[g.s];
        g.uoff = h->uplan[g.s];
    }
    // triangular solves
    // algorithmic bytes of the solves: each stored factor entry is read once
    // per direction (element size of the factor type)
    const double wsz = h->use_fp32 ? 4.0 : 8.0;
    auto tri_bytes = [&](int32_t s, bool with_l21) {
        double n = ncol(s), c = S.snrow[s];
        return wsz * (n * (n + 1) / 2 + (with_l21 ? (c - n) * n : 0.0));
    };
    // each level's list of supernodes that are not wide: the tiny ones (one
    // wave each) first; wide supernodes go to per-block tasks instead
    auto level_list = [&](int32_t l, std::vector<int32_t>& out) {
        int32_t ntiny = 0;
        for (int pass = 0; pass < 2; ++pass)
            for (int32_t t = S.levelptr[l]; t < S.levelptr[l + 1]; ++t) {
                int32_t s = S.levelsn[t];
                if (wide(s)) continue;
                bool tiny = S.snrow[s] <= kTinyRows;
                if (tiny != (pass == 0)) continue;
                out.push_back(s);
                ntiny += tiny;
            }
        return ntiny;
    };
    // bytes of the inverse-based triangle of a wide supernode: the stored
    // off-diagonal inverse blocks plus the diagonal-block inverses
    auto xinv_bytes = [&](int32_t s) {
        double nb = nblk_of(s), n = ncol(s);
        return wsz * (nb * (nb - 1) / 2 * kTile * kTile + n * kTile);
    };
    // lanes per tiny supernode for a level of tiny supernodes only: 32 (two
    // per wave) when every one has at most 32 rows, else 64
    auto tiny_lanes = [&](const std::vector<int32_t>& lst, int32_t off, int32_t count, int32_t ntiny,
                          int32_t nx) {
        if (ntiny != count || nx != 0) return 64;
        for (int32_t t = off; t < off + count; ++t)
            if (S.snrow[lst[t]] > 32) return 64;
        return 32;
    };
    for (int32_t l = 0; l < S.nlevels; ++l) {
        int32_t off = (int32_t)h->list_fwd.size(), to = (int32_t)h->fgemv.size(), xo = (int32_t)h->fxs.size();
        double w = 0, wg = 0;
        int32_t ntiny = level_list(l, h->list_fwd);
        for (int32_t t = S.levelptr[l]; t < S.levelptr[l + 1]; ++t) {
            int32_t s = S.levelsn[t];
            int32_t nc = S.sfirst[s + 1] - S.sfirst[s], c = S.snrow[s];
            bool big = is_big_solve(c, nc) || wide(s);
            if (wide(s)) {
                w += xinv_bytes(s);
                for (int32_t i = 0; i < nblk_of(s); ++i) h->fxs.push_back({s, i});
            } else {
                w += tri_bytes(s, !big);
            }
            if (big) {
                h->sbig[s] = wide(s) ? 2 : 1;
                wg += wsz * (double)(c - nc) * nc;
                for (int32_t b = 0; nc + b * 64 < c; ++b) h->fgemv.push_back({s, b});
            }
        }
        Op o{OP_FWD, off, (int32_t)h->list_fwd.size() - off, w, 0, ntiny};
        o.xoff = xo;
        o.xcount = (int32_t)h->fxs.size() - xo;
        o.wave = tiny_lanes(h->list_fwd, off, o.count, ntiny, o.xcount);
        h->sops_fwd.push_back(o);
        if ((int32_t)h->fgemv.size() > to) h->sops_fwd.push_back({OP_FWD_GEMV, to, (int32_t)h->fgemv.size() - to, wg});
    }
    for (int32_t l = S.nlevels - 1; l >= 0; --l) {
        int32_t off = (int32_t)h->list_bwd.size(), to = (int32_t)h->bgemv.size(), xo = (int32_t)h->bxs.size();
        double w = 0, wg = 0;
        int32_t ntiny = level_list(l, h->list_bwd);
        for (int32_t t = S.levelptr[l]; t < S.levelptr[l + 1]; ++t) {
            int32_t s = S.levelsn[t];
            int32_t nc = S.sfirst[s + 1] - S.sfirst[s], c = S.snrow[s];
            if (wide(s)) {
                w += xinv_bytes(s);
                for (int32_t j = 0; j < nblk_of(s); ++j) h->bxs.push_back({s, j});
            } else {
                w += tri_bytes(s, !h->sbig[s]);
            }
            if (h->sbig[s]) {
                wg += wsz * (double)(c - nc) * nc;
                for (int32_t b = 0; b * 16 < nc; ++b) h->bgemv.push_back({s, b});
            }
        }
        if ((int32_t)h->bgemv.size() > to) h->sops_bwd.push_back({OP_BWD_GEMV, to, (int32_t)h->bgemv.size() - to, wg});
        Op o{OP_BWD, off, (int32_t)h->list_bwd.size() - off, w, 0, ntiny};
        o.xoff = xo;
        o.xcount = (int32_t)h->bxs.size() - xo;
        o.wave = tiny_lanes(h->list_bwd, off, o.count, ntiny, o.xcount);
        h->sops_bwd.push_back(o);
    }
    // SpMV-class kernels: 12 bytes per nonzero (value + index) plus vectors
    const double nnz = (double)h->nnz, m = h->m, n = h->n;
    h->work_adat = 12.0 * nnz + 8.0 * n + 12.0 * (double)h->P.colind.size();
    h->work_rhs = 12.0 * nnz + 8.0 * (m + 1) + 16.0 * n + 8.0 * m + 4.0 * m;
    h->work_backsub = 12.0 * nnz + 8.0 * (n + 1) + 8.0 * m + 24.0 * n;
}

// stream, events, A in CSC and CSR, the boundary and work vectors
void setup_device_common(help_gpu* h) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw HipError("no HIP device visible", HELP_ERR_NOGPU);
    HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->st_side, hipStreamNonBlocking));
    for (auto& e : h->ev) HIPCHK(hipEventCreate(&e));
    for (auto& e : h->ev_side) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&h->h_pin, 8 * sizeof(double), hipHostMallocDefault));
    hipStream_t st = h->st;
    h->d_colptr.upload(h->A.colptr, st);
    h->d_rowind.upload(h->A.rowind, st);
    h->d_val.upload(h->A.val, st);
    h->d_rowptr.upload(h->R.rowptr, st);
    h->d_colind.upload(h->csrcol, st);
    h->d_csrval.upload(h->csrval, st);
    const int32_t m = h->m, n = h->n;
    // b1, x1, b2, x2, rp, t and c hold two right-hand sides (help_gpu_solve2)
    h->d_d.alloc(n); h->d_dinv.alloc(n); h->d_b1.alloc(2 * (size_t)n); h->d_x1.alloc(2 * (size_t)n);
    h->d_tn.alloc(n);
    h->d_b2.alloc(2 * (size_t)m); h->d_x2.alloc(2 * (size_t)m); h->d_rp.alloc(2 * (size_t)m);
    h->d_t.alloc(2 * (size_t)m); h->d_r.alloc(2 * (size_t)m); h->d_c.alloc(2 * (size_t)m);
    h->d_z.alloc(2 * (size_t)m); h->d_p.alloc(2 * (size_t)m); h->d_q.alloc(2 * (size_t)m);
    h->d_part.alloc(kDotPartials); h->d_scal.alloc(8);
    h->d_err.alloc(1);
    std::vector<help::RowChunk> chunks;
    std::vector<int32_t> hrow, hptr, mrow;
    help::build_row_split(m, h->R.rowptr, chunks, hrow, hptr, mrow);
    for (int32_t i = 0; i < m; ++i)
        h->max_row_nnz = std::max<int32_t>(h->max_row_nnz, (int32_t)(h->R.rowptr[i + 1] - h->R.rowptr[i]));
    h->d_chunks.upload(chunks, st);
    h->d_hrow.upload(hrow, st);
    h->d_hptr.upload(hptr, st);
    h->d_mrow.upload(mrow, st);
    h->d_rpart.alloc(std::max<size_t>(1, chunks.size()));
    h->rows.m = m;
    h->rows.nchunk = (int32_t)chunks.size();
    h->rows.nheavy = (int32_t)hrow.size();
    h->rows.chunks = h->d_chunks.p;
    h->rows.hrow = h->d_hrow.p;
    h->rows.hptr = h->d_hptr.p;
    h->rows.part = h->d_rpart.p;
    h->rows.nmed = (int32_t)mrow.size();
    h->rows.mrow = h->d_mrow.p;
}

// symbolic structures, schedules and factor storage of the direct modes
void ensure_fp64_factor(help_gpu* h);

void setup_device_factor(help_gpu* h) {
    hipStream_t st = h->st;
    const Supernodal& S = h->S;
    h->d_pairptr.upload(h->P.pairptr, st);
    h->d_wnz.alloc(std::max<int64_t>(1, h->nnz));
    h->d_pair_p.upload(h->P.pair_p, st);
    h->d_pair_q.upload(h->P.pair_q, st);
    h->d_colof.upload(h->colof, st);
    h->d_dest.upload(h->slot_dest, st);
    h->d_diagpos.upload(h->diagpos, st);
    h->d_iperm.upload(S.iperm, st);
    h->d_sfirst.upload(S.sfirst, st);
    h->d_snrow.upload(S.snrow, st);
    h->d_loff.upload(S.loff, st);
    h->d_uoff.upload(h->uplan, st);
    h->d_rowoff.upload(S.rowoff, st);
    h->d_rowidx.upload(S.rowidx, st);
    h->d_reloff.upload(S.reloff, st);
    h->d_relmap.upload(S.relmap, st);
    h->d_childptr.upload(S.childptr, st);
    h->d_child.upload(S.child, st);
    h->d_dinvoff.upload(h->dinvoff, st);
    h->d_list_small.upload(h->list_small, st);
    h->d_list_fwd.upload(h->list_fwd, st);
    h->d_list_bwd.upload(h->list_bwd, st);
    h->d_big.upload(h->sbig, st);
    h->d_ea.upload(h->ea, st);
    h->d_earef.upload(h->earef, st);
    h->d_potrf.upload(h->potrf, st);
    h->d_panel.upload(h->panel, st);
    h->d_pcopy.upload(h->pcopy, st);
    h->d_trsm.upload(h->trsm, st);
    h->d_gemm.upload(h->gemm, st);
    h->d_fgemv.upload(h->fgemv, st);
    h->d_bgemv.upload(h->bgemv, st);
    h->d_fxs.upload(h->fxs, st);
    h->d_bxs.upload(h->bxs, st);
    h->d_xtask.upload(h->xtask, st);
    h->d_xoff.upload(h->xoff, st);
    // two right-hand sides' worth of solve vectors (help_gpu_solve2)
    h->w_len = std::max<int64_t>(1, S.reloff[S.nsuper]);
    h->d_y.alloc(2 * (size_t)h->m);
    h->d_W.alloc(2 * (size_t)h->w_len);
    int64_t nl = S.loff[S.nsuper];
    int64_t nd = h->dinvoff[S.nsuper];
    // update matrices: shared by both precisions (at most one factorization
    // runs at a time and nothing reads them after it)
    h->d_Ubuf.alloc(std::max<int64_t>(1, h->uplan[S.nsuper]));
    if (h->use_uacc) h->d_Uaccbuf.alloc(std::max<int64_t>(1, h->uplan[S.nsuper]));
    if (h->use_fp32) {
        h->d_slot_row.upload(h->slot_row, st);
        h->d_slot_col.upload(h->P.colind, st);
        h->d_scale.alloc(h->m);
        h->d_L32.alloc(std::max<int64_t>(1, nl));
        h->d_Dinv32.alloc(std::max<int64_t>(1, nd));
        if (h->use_xinv) {
            h->d_X32.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
            h->d_S32.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
        }
        // the FP64 factor the run may switch to (refinement cost, or a
        // refinement failure) is allocated here, not in the middle of a run
        ensure_fp64_factor(h);
    } else {
        h->d_L64.alloc(std::max<int64_t>(1, nl));
        h->d_Dinv64.alloc(std::max<int64_t>(1, nd));
        if (!h->pcopy.empty()) h->d_Lstage64.alloc(std::max<int64_t>(1, nd));
        if (h->use_xinv) {
            h->d_X64.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
            h->d_S64.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
        }
    }
    DevSym& y = h->sym;
    y.nsuper = S.nsuper;
    y.sfirst = h->d_sfirst.p; y.snrow = h->d_snrow.p; y.loff = h->d_loff.p; y.uoff = h->d_uoff.p;
    y.rowoff = h->d_rowoff.p; y.rowidx = h->d_rowidx.p; y.reloff = h->d_reloff.p;
    y.relmap = h->d_relmap.p; y.childptr = h->d_childptr.p; y.child = h->d_child.p;
    y.dinvoff = h->d_dinvoff.p;
    y.xoff = h->d_xoff.p;
}

void setup_device(help_gpu* h) {
    setup_device_common(h);
    if (h->use_pcg) help::pcg_setup(h);
    else setup_device_factor(h);
    HIPCHK(hipStreamSynchronize(h->st));
    h->device = true;
}

// inverse-block storage of the factor precision T, and its partial sums
template <typename T>
T* xinv_store(help_gpu* h) {
    if (std::is_same<T, float>::value) return (T*)h->d_X32.p;
    return (T*)h->d_X64.p;
}
template <typename T>
T* uacc_store(help_gpu* h) {
    return (T*)h->d_Uaccbuf.p;
}

template <typename T>
T* xinv_partial(help_gpu* h) {
    if (std::is_same<T, float>::value) return (T*)h->d_S32.p;
    return (T*)h->d_S64.p;
}

template <typename T>
void run_factor(help_gpu* h, T* L, T* U, T* Dinv) {
    hipStream_t st = h->st;
    h->prof_fp32 = std::is_same<T, float>::value;
    // The inverse blocks of a level need only that level's finished L11 and
    // Dinv blocks, and nothing on the main stream reads them before the
    // solves: each run of XINV launches goes to the side stream behind a
    // fork event, and the main stream joins the side stream at the end.
    bool forked = false, side_used = false;
    for (const Op& o : h->fops) {
        const bool side = o.kind == OP_XINV;
        if (side && !forked) {
            HIPCHK(hipEventRecord(h->ev_side[0], st));
            HIPCHK(hipStreamWaitEvent(h->st_side, h->ev_side[0], 0));
            side_used = true;
        }
        forked = side;
        const bool fused = std::is_same<T, double>::value &&
                           ((o.kind == OP_POTRF && o.xcount > 0) || (o.kind == OP_TRSM && o.fused));
        if (fused && o.kind == OP_TRSM) continue;   // done by the fused panel launch
        ProfScope ps(h, o.kind, fused ? o.work2 : o.work);
        switch (o.kind) {
            case OP_SMALL:
                if (o.wave)
                    launch_small_wave<T>(h->d_list_small.p + o.off, o.count, h->sym, L, U, Dinv, h->d_err.p, o.lds,
                                         o.wave, st);
                else
                    launch_small_fronts<T>(h->d_list_small.p + o.off, o.count, h->sym, L, U, Dinv, h->d_err.p, o.lds, st);
                break;
            case OP_EA: launch_extend_add<T>(h->d_ea.p + o.off, o.count, h->d_earef.p, h->sym, L, U, o.depth, st); break;
            case OP_POTRF:
                if constexpr (std::is_same<T, double>::value) {
                    if (fused) {
                        launch_panel_fp64(h->d_panel.p + o.xoff, o.xcount, h->sym, L, Dinv, h->d_Lstage64.p,
                                          h->d_err.p, st);
                        break;
                    }
                }
                launch_potrf_diag<T>(h->d_potrf.p + o.off, o.count, h->sym, L, Dinv, h->d_err.p, st);
                break;
            case OP_TRSM: launch_trsm_rows<T>(h->d_trsm.p + o.off, o.count, h->sym, L, Dinv, st); break;
            case OP_GEMM: launch_gemm<T>(h->d_gemm.p + o.off, o.count, h->sym, L, U, uacc_store<T>(h), st); break;
            case OP_XINV: launch_xinv<T>(h->d_xtask.p + o.off, o.count, h->sym, L, Dinv, xinv_store<T>(h), xinv_partial<T>(h), h->st_side); break;
            default: break;
        }
    }
    if constexpr (std::is_same<T, double>::value) {
        // the fused panels' L11 blocks into L (nothing reads them from L
        // during the factorization)
        if (!h->pcopy.empty())
            launch_copy_diag_fp64(h->d_pcopy.p, (int32_t)h->pcopy.size(), h->sym, h->d_Lstage64.p, L, st);
    }
    if (side_used) {
        HIPCHK(hipEventRecord(h->ev_side[1], h->st_side));
        HIPCHK(hipStreamWaitEvent(st, h->ev_side[1], 0));
    }
    h->prof_fp32 = false;
    HIPCHK(hipGetLastError());
}

// x (permuted numbering, FP64) <- (L L^T)^-1 x; the forward result goes
// to d_y, the backward one back to x.  nrhs right-hand sides at x + r m
// (r < nrhs) are solved by the same launches, each with exactly the
// operations of a single solve.
template <typename T>
void run_trisolve(help_gpu* h, const T* L, const T* Dinv, double* x, int nrhs = 1) {
    hipStream_t st = h->st;
    h->prof_fp32 = std::is_same<T, float>::value;
    const T* X = xinv_store<T>(h);
    double* y = h->d_y.p;
    RhsStride rs;
    rs.v = h->m;
    rs.w = h->w_len;
    for (const Op& o : h->sops_fwd) {
        ProfScope ps(h, o.kind, o.work * nrhs);
        if (o.kind == OP_FWD)
            launch_fwd_level<T>(h->d_list_fwd.p + o.off, o.count, o.ntiny, h->d_fxs.p + o.xoff, o.xcount, h->sym,
                                h->d_big.p, L, Dinv, X, x, y, h->d_W.p, nrhs, rs, st, o.wave);
        else
            launch_fwd_gemv<T>(h->d_fgemv.p + o.off, o.count, h->sym, L, y, h->d_W.p, nrhs, rs, st);
    }
    for (const Op& o : h->sops_bwd) {
        ProfScope ps(h, o.kind, o.work * nrhs);
        if (o.kind == OP_BWD)
            launch_bwd_level<T>(h->d_list_bwd.p + o.off, o.count, o.ntiny, h->d_bxs.p + o.xoff, o.xcount, h->sym,
                                h->d_big.p, L, Dinv, X, y, h->d_t.p, x, nrhs, rs, st, o.wave);
        else
            launch_bwd_gemv<T>(h->d_bgemv.p + o.off, o.count, h->sym, L, x, h->d_t.p, nrhs, rs, st);
    }
    h->prof_fp32 = false;
    HIPCHK(hipGetLastError());
}

double dot_dev(help_gpu* h, const double* a, const double* b, int64_t n) {
    launch_dot(a, b, n, h->d_part.p, h->d_scal.p, h->st);
    double v = 0;
    HIPCHK(hipMemcpyAsync(&v, h->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return v;
}

// z = P (L L^T)^-1 P^T r with the factor of type T (vectors FP64)
template <typename T>
void apply_factor(help_gpu* h, const double* r, double* z) {
    const T* Lf = std::is_same<T, float>::value ? (const T*)h->d_L32.p : (const T*)h->d_L64.p;
    const T* Df = std::is_same<T, float>::value ? (const T*)h->d_Dinv32.p : (const T*)h->d_Dinv64.p;
    launch_scatter_perm(h->m, r, h->d_iperm.p, h->d_rp.p, h->st);
    if (std::is_same<T, float>::value) launch_scale_perm(h->m, h->d_rp.p, h->d_scale.p, h->d_iperm.p, h->st);
    run_trisolve<T>(h, Lf, Df, h->d_rp.p);
    if (std::is_same<T, float>::value) launch_scale_perm(h->m, h->d_rp.p, h->d_scale.p, h->d_iperm.p, h->st);
    launch_gather_perm(h->m, h->d_rp.p, h->d_iperm.p, z, h->st);
}

// Preconditioned CG in FP64 on M = A D^-1 A^T + beta I with the factor of
// type T as preconditioner, started from the factor's solution.  On entry
// d_rp holds the permuted right-hand side; on exit d_x2 holds the solution
// (original order).  Stops at ||r|| <= refine_tol ||rhs||, after max_it
// iterations, or (FP64 factor) when an iteration fails to halve the
// residual: the rounding floor; returns iterations.  The final relative
// residual is measured in FP64 also when max_it is 0.
template <typename T>
int pcg_factor(help_gpu* h, double& rel, int max_it) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    double *rhs = h->d_c.p, *r = h->d_r.p, *x = h->d_x2.p, *z = h->d_z.p, *p = h->d_p.p, *q = h->d_q.p;
    const T* Lf = std::is_same<T, float>::value ? (const T*)h->d_L32.p : (const T*)h->d_L64.p;
    const T* Df = std::is_same<T, float>::value ? (const T*)h->d_Dinv32.p : (const T*)h->d_Dinv64.p;
    launch_gather_perm(m, h->d_rp.p, h->d_iperm.p, rhs, st);
    if (std::is_same<T, float>::value) launch_scale_perm(m, h->d_rp.p, h->d_scale.p, h->d_iperm.p, st);
    run_trisolve<T>(h, Lf, Df, h->d_rp.p);
    if (std::is_same<T, float>::value) launch_scale_perm(m, h->d_rp.p, h->d_scale.p, h->d_iperm.p, st);
    launch_gather_perm(m, h->d_rp.p, h->d_iperm.p, x, st);
    launch_dinv(h->d_d.p, h->d_dinv.p, n, st);
    auto matvec = [&](const double* v, double* out) {
        launch_normal_matvec(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_colptr.p,
                             h->d_rowind.p, h->d_val.p, h->d_dinv.p, h->opts.beta, v, h->d_tn.p, out, st);
    };
    matvec(x, q);
    launch_sub(m, rhs, q, r, st);
    if (max_it <= 0) {
        // no refinement: the residual is only reported, so its two norms are
        // copied to pinned memory and read after the solve's final
        // synchronization (no host round trip inside the solve)
        launch_dot(rhs, rhs, m, h->d_part.p, h->d_scal.p + 2, st);
        launch_dot(r, r, m, h->d_part.p, h->d_scal.p + 3, st);
        HIPCHK(hipMemcpyAsync(h->h_pin, h->d_scal.p + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        h->rel_deferred = true;
        rel = 0;
        return 0;
    }
    const double rr0 = dot_dev(h, rhs, rhs, m);
    if (rr0 == 0.0) { rel = 0; return 0; }
    double rr = dot_dev(h, r, r, m);
    rel = std::sqrt(rr / rr0);
    if (rel <= h->opts.refine_tol || max_it <= 0) return 0;
    apply_factor<T>(h, r, z);
    HIPCHK(hipMemcpyAsync(p, z, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    double rz = dot_dev(h, r, z, m);
    int it = 0;
    while (it < max_it && rz > 0) {
        ++it;
        matvec(p, q);
        double pq = dot_dev(h, p, q, m);
        if (!(pq > 0)) break;
        double alpha = rz / pq;
        launch_axpy(m, alpha, p, x, st);
        launch_axpy(m, -alpha, q, r, st);
        const double prev = rel;
        rr = dot_dev(h, r, r, m);
        rel = std::sqrt(rr / rr0);
        if (rel <= h->opts.refine_tol) break;
        // with the FP64 factor one iteration reaches the rounding floor; the
        // CG residual norm is not monotone, so the FP32 factor runs on
        if (std::is_same<T, double>::value && rel > 0.5 * prev) break;
        apply_factor<T>(h, r, z);
        double rz_new = dot_dev(h, r, z, m);
        double beta = rz_new / rz;
        rz = rz_new;
        launch_xpby(m, z, beta, p, st);
    }
    return it;
}

// pcg_factor for two right-hand sides at once (d_rp slots 0 and 1 on entry,
// d_x2 slots 0 and 1 on exit): the two refinements run in lockstep and share
// the factor's triangular solves (one batched solve per step while both are
// running); every vector operation, dot and scalar test of a system is the
// one pcg_factor makes, so each solution is the same bit for bit.
template <typename T>
void pcg_factor2(help_gpu* h, double rel[2], int it[2], int max_it) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    const T* Lf = std::is_same<T, float>::value ? (const T*)h->d_L32.p : (const T*)h->d_L64.p;
    const T* Df = std::is_same<T, float>::value ? (const T*)h->d_Dinv32.p : (const T*)h->d_Dinv64.p;
    const bool f32 = std::is_same<T, float>::value;
    double* rp = h->d_rp.p;
    auto slot = [&](double* base, int r) { return base + (size_t)r * m; };
    for (int r = 0; r < 2; ++r) {
        launch_gather_perm(m, slot(rp, r), h->d_iperm.p, slot(h->d_c.p, r), st);
        if (f32) launch_scale_perm(m, slot(rp, r), h->d_scale.p, h->d_iperm.p, st);
    }
    run_trisolve<T>(h, Lf, Df, rp, 2);
    for (int r = 0; r < 2; ++r) {
        if (f32) launch_scale_perm(m, slot(rp, r), h->d_scale.p, h->d_iperm.p, st);
        launch_gather_perm(m, slot(rp, r), h->d_iperm.p, slot(h->d_x2.p, r), st);
    }
    launch_dinv(h->d_d.p, h->d_dinv.p, n, st);
    auto matvec = [&](const double* v, double* out) {
        launch_normal_matvec(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_colptr.p,
                             h->d_rowind.p, h->d_val.p, h->d_dinv.p, h->opts.beta, v, h->d_tn.p, out, st);
    };
    // z_r = P (L L^T)^-1 P^T r_r for the systems in `act` (slots of d_rp in
    // that order), one batched triangular solve
    auto apply2 = [&](const int* act, int na) {
        for (int a = 0; a < na; ++a) {
            launch_scatter_perm(m, slot(h->d_r.p, act[a]), h->d_iperm.p, slot(rp, a), st);
            if (f32) launch_scale_perm(m, slot(rp, a), h->d_scale.p, h->d_iperm.p, st);
        }
        run_trisolve<T>(h, Lf, Df, rp, na);
        for (int a = 0; a < na; ++a) {
            if (f32) launch_scale_perm(m, slot(rp, a), h->d_scale.p, h->d_iperm.p, st);
            launch_gather_perm(m, slot(rp, a), h->d_iperm.p, slot(h->d_z.p, act[a]), st);
        }
    };
    double rr0[2], rz[2];
    bool run[2];
    for (int r = 0; r < 2; ++r) {
        matvec(slot(h->d_x2.p, r), slot(h->d_q.p, r));
        launch_sub(m, slot(h->d_c.p, r), slot(h->d_q.p, r), slot(h->d_r.p, r), st);
    }
    for (int r = 0; r < 2; ++r) {
        it[r] = 0;
        rel[r] = 0;
        run[r] = false;
        rr0[r] = dot_dev(h, slot(h->d_c.p, r), slot(h->d_c.p, r), m);
        if (rr0[r] == 0.0) continue;
        const double rr = dot_dev(h, slot(h->d_r.p, r), slot(h->d_r.p, r), m);
        rel[r] = std::sqrt(rr / rr0[r]);
        run[r] = !(rel[r] <= h->opts.refine_tol || max_it <= 0);
    }
    int act[2], na = 0;
    for (int r = 0; r < 2; ++r) if (run[r]) act[na++] = r;
    if (na == 0) return;
    apply2(act, na);
    for (int a = 0; a < na; ++a) {
        const int r = act[a];
        HIPCHK(hipMemcpyAsync(slot(h->d_p.p, r), slot(h->d_z.p, r), sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        rz[r] = dot_dev(h, slot(h->d_r.p, r), slot(h->d_z.p, r), m);
    }
    while (true) {
        na = 0;
        for (int r = 0; r < 2; ++r) {
            if (!run[r]) continue;
            if (!(it[r] < max_it && rz[r] > 0)) { run[r] = false; continue; }
            ++it[r];
            double* p = slot(h->d_p.p, r);
            double* q = slot(h->d_q.p, r);
            double* rv = slot(h->d_r.p, r);
            matvec(p, q);
            const double pq = dot_dev(h, p, q, m);
            if (!(pq > 0)) { run[r] = false; continue; }
            const double alpha = rz[r] / pq;
            launch_axpy(m, alpha, p, slot(h->d_x2.p, r), st);
            launch_axpy(m, -alpha, q, rv, st);
            const double prev = rel[r];
            const double rr = dot_dev(h, rv, rv, m);
            rel[r] = std::sqrt(rr / rr0[r]);
            if (rel[r] <= h->opts.refine_tol) { run[r] = false; continue; }
            if (std::is_same<T, double>::value && rel[r] > 0.5 * prev) { run[r] = false; continue; }
            act[na++] = r;
        }
        if (na == 0) break;
        apply2(act, na);
        for (int a = 0; a < na; ++a) {
            const int r = act[a];
            const double rz_new = dot_dev(h, slot(h->d_r.p, r), slot(h->d_z.p, r), m);
            const double beta = rz_new / rz[r];
            rz[r] = rz_new;
            launch_xpby(m, slot(h->d_z.p, r), beta, slot(h->d_p.p, r), st);
        }
    }
}

// FP64 factor storage, allocated on first use in the FP32 mode (after a
// refinement failure)
void ensure_fp64_factor(help_gpu* h) {
    if (h->d_L64.p) return;
    const Supernodal& S = h->S;
    h->d_L64.alloc(std::max<int64_t>(1, S.loff[S.nsuper]));
    h->d_Dinv64.alloc(std::max<int64_t>(1, h->dinvoff[S.nsuper]));
    if (!h->pcopy.empty()) h->d_Lstage64.alloc(std::max<int64_t>(1, h->dinvoff[S.nsuper]));
    if (h->use_xinv) {
        h->d_X64.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
        h->d_S64.alloc(std::max<int64_t>(1, h->xoff[S.nsuper]));
    }
}

// ADA^T (+ beta on the diagonal) written into the zeroed factor storage of
// the requested precision; d_dinv holds 1/d.  The update matrices are not
// cleared here: the factorization's extend-add and small-front kernels zero
// each front's update matrix before the first contribution to it.
void assemble(help_gpu* h, bool f32) {
    hipStream_t st = h->st;
    const int32_t m = h->m;
    const Supernodal& S = h->S;
    int64_t nl = S.loff[S.nsuper];
    int64_t nslot = (int64_t)h->P.colind.size();
    if (f32) {
        HIPCHK(hipMemsetAsync(h->d_L32.p, 0, sizeof(float) * nl, st));
        // factor S (A D^-1 A^T + beta I) S, S = diag(1/sqrt(diag)): unit
        // diagonal, entries bounded by 1, no FP32 overflow
        launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_dinv.p, h->d_t.p, true, st);
        launch_diag_scale(m, h->d_t.p, h->opts.beta, h->d_scale.p, st);
        {
            ProfScope ps(h, HELP_K_ADAT, h->work_adat);
            launch_adat_scaled_f32(h->d_pairptr.p, h->d_pair_p.p, h->d_pair_q.p, h->d_colof.p, h->d_val.p,
                                   h->d_dinv.p, h->nnz, h->d_wnz.p, nslot, h->d_slot_row.p, h->d_slot_col.p, h->d_scale.p,
                                   h->d_dest.p, h->d_L32.p, st);
        }
        launch_add_beta_scaled_f32(h->d_diagpos.p, m, h->opts.beta, h->d_scale.p, h->d_L32.p, st);
    } else {
        HIPCHK(hipMemsetAsync(h->d_L64.p, 0, sizeof(double) * nl, st));
        {
            ProfScope ps(h, HELP_K_ADAT, h->work_adat);
            launch_adat_assemble(h->d_pairptr.p, h->d_pair_p.p, h->d_pair_q.p, h->d_colof.p, h->d_val.p,
                                 h->d_dinv.p, h->nnz, h->d_wnz.p, nslot, h->d_dest.p, h->d_L64.p, nullptr, st);
        }
        launch_add_beta(h->d_diagpos.p, m, h->opts.beta, h->d_L64.p, nullptr, st);
    }
}

// The factorization's launches depend only on the symbolic schedule and the
// handle's buffers, so with opts.factor_graph they are captured once per
// precision (both streams: the side stream joins the capture through the
// fork event and rejoins the main stream before it ends) and replayed on
// every later renew.  The capture records kernels and events only (the
// renew's memsets stay outside).  Not used while profiling (the profiling
// events are per launch).
template <typename T>
void factor_launches(help_gpu* h) {
    if constexpr (std::is_same<T, float>::value)
        run_factor<float>(h, h->d_L32.p, reinterpret_cast<float*>(h->d_Ubuf.p), h->d_Dinv32.p);
    else
        run_factor<double>(h, h->d_L64.p, h->d_Ubuf.p, h->d_Dinv64.p);
}

void factor(help_gpu* h, bool f32) {
    if (h->opts.factor_graph && !h->prof) {
        hipGraphExec_t& ge = h->fgraph[f32 ? 1 : 0];
        if (!ge) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
            try {
                if (f32) factor_launches<float>(h);
                else factor_launches<double>(h);
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(h->st, &dead);
                if (dead) (void)hipGraphDestroy(dead);
                throw;
            }
            HIPCHK(hipStreamEndCapture(h->st, &g));
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { ge = nullptr; HIPCHK(e); }
        }
        HIPCHK(hipGraphLaunch(ge, h->st));
    } else if (f32) {
        factor_launches<float>(h);
    } else {
        factor_launches<double>(h);
    }
    h->factor_fp32 = f32;
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

int fail(help_gpu* h, const std::exception& e) {
    h->err = e.what();
    if (auto* he = dynamic_cast<const HipError*>(&e)) return he->code;
    return HELP_ERR_ARG;
}

}  // namespace

extern "C" {

void help_opts_default(help_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->mode = HELP_MODE_DIRECT_FP64;
    o->beta = 1e-5;
    o->refine_max = 20;
    o->refine_tol = 1e-12;
    o->refine_switch = 2.5;
    o->pcg_k = 500;
    o->pcg_max_iter = 5000;
    o->pcg_tol = 1e-10;
    o->pcg_fp32_precond = 1;
}

int help_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* help_last_create_error(void) { return g_create_error.c_str(); }

int help_gpu_create(int32_t m, int32_t n, int64_t nnz, const int64_t* colptr,
                    const int32_t* rowind, const double* val, const help_opts* opts,
                    help_gpu** out) {
    g_create_error.clear();
    if (!out || m <= 0 || n <= 0 || nnz < 0 || !colptr || (nnz > 0 && (!rowind || !val))) {
        g_create_error = "help_gpu_create: bad arguments";
        return HELP_ERR_ARG;
    }
    *out = nullptr;
    help_gpu* h = new (std::nothrow) help_gpu();
    if (!h) return HELP_ERR_NOMEM;
    try {
        if (opts) h->opts = *opts; else help_opts_default(&h->opts);
        if (colptr[0] != 0 || colptr[n] != nnz) throw std::runtime_error("colptr inconsistent with nnz");
        for (int32_t j = 0; j < n; ++j) {
            if (colptr[j + 1] < colptr[j]) throw std::runtime_error("colptr not monotone");
            for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
                if (rowind[p] < 0 || rowind[p] >= m) throw std::runtime_error("row index out of range");
                if (p > colptr[j] && rowind[p] <= rowind[p - 1])
                    throw std::runtime_error("row indices not strictly ascending within a column");
            }
        }
        if (nnz > INT32_MAX) throw std::runtime_error("nnz exceeds the int32 index range");
        h->m = m; h->n = n; h->nnz = nnz;
        h->use_fp32 = h->opts.mode == HELP_MODE_DIRECT_FP32;
        h->use_pcg = h->opts.mode == HELP_MODE_PCG || h->opts.mode == HELP_MODE_PCG_LEARNED;
        h->learned = h->opts.mode == HELP_MODE_PCG_LEARNED;
        if (h->opts.mode < HELP_MODE_DIRECT_FP64 || h->opts.mode > HELP_MODE_PCG_LEARNED)
            throw std::runtime_error("unknown mode");
        h->A.m = m; h->A.n = n;
        h->A.colptr.assign(colptr, colptr + n + 1);
        h->A.rowind.assign(rowind, rowind + nnz);
        h->A.val.assign(val, val + nnz);
        // phase timing for verbose runs (stderr)
        auto t_last = std::chrono::steady_clock::now();
        auto phase = [&](const char* what) {
            if (!h->opts.verbose) return;
            const auto now = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[help] create: %-18s %8.3f s\n", what,
                         std::chrono::duration<double>(now - t_last).count());
            t_last = now;
        };
        h->R = csc_to_csr(h->A);
        phase("csr");
        if (h->use_pcg) {
            // matrix-free: no ADA^T pattern, ordering or factor
            h->csrcol = h->R.colind;
            h->csrval.resize(h->nnz);
            for (int64_t t = 0; t < h->nnz; ++t) h->csrval[t] = h->A.val[h->R.cscpos[t]];
            const double nnz_d = (double)nnz, md = m, nd = n;
            h->work_rhs = 12.0 * nnz_d + 16.0 * nd + 16.0 * md;
            h->work_backsub = 12.0 * nnz_d + 8.0 * (nd + 1) + 8.0 * md + 24.0 * nd;
        } else {
            // The ADA^T pattern with its contribution pairs is needed only by
            // the assembly maps: it is built on a thread of its own while the
            // ordering and the symbolic analysis run (the ordering reads the
            // adjacency, which comes from A directly).  The schedules, which
            // need only the analysis, are built beside the assembly maps.
            std::exception_ptr side_err;
            std::thread side([&] {
                try { h->P = adat_pattern(h->A, h->R, true); } catch (...) { side_err = std::current_exception(); }
            });
            struct Joiner {
                std::thread& t;
                ~Joiner() { if (t.joinable()) t.join(); }
            } join_side{side};
            std::vector<int64_t> adjptr;
            std::vector<int32_t> adj;
            adat_adjacency(h->A, h->R, adjptr, adj);
            phase("adjacency");
            RelaxParams rp;
            std::vector<int32_t> perm;
            if (h->opts.ordering == HELP_ORDER_ND) perm = help::nd_order(m, adjptr, adj);
            else if (h->opts.ordering != HELP_ORDER_AMD) throw std::runtime_error("unknown ordering");
            phase("ordering (ND)");
            h->S = analyze(m, adjptr, adj, perm, rp);
            phase("analyze");
            side.join();
            if (side_err) std::rethrow_exception(side_err);
            phase("adat pattern (rest)");
            // (the FP32 mode's row of every ADA^T slot, for the scaled
            // assembly, is filled on the same thread as the schedules)
            std::thread sched([&] {
                try {
                    build_schedules(h);
                    if (h->use_fp32) {
                        h->slot_row.resize(h->P.colind.size());
                        for (int32_t i = 0; i < m; ++i)
                            for (int64_t t = h->P.rowptr[i]; t < h->P.rowptr[i + 1]; ++t) h->slot_row[t] = i;
                    }
                } catch (...) { side_err = std::current_exception(); }
            });
            struct Joiner join_sched{sched};
            build_assembly_maps(h);
            phase("assembly maps");
            sched.join();
            if (side_err) std::rethrow_exception(side_err);
            phase("schedules (rest)");
        }
        if (!h->opts.symbolic_only) setup_device(h);
        phase("device setup");
        h->prof = h->opts.profile != 0;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        int code = HELP_ERR_ARG;
        if (auto* he = dynamic_cast<const HipError*>(&e)) code = he->code;
        if (auto* ba = dynamic_cast<const std::bad_alloc*>(&e)) { (void)ba; code = HELP_ERR_NOMEM; }
        delete h;
        return code;
    }
    *out = h;
    return HELP_OK;
}

int help_gpu_destroy(help_gpu* h) {
    delete h;
    return HELP_OK;
}

const char* help_gpu_last_error(const help_gpu* h) { return h ? h->err.c_str() : ""; }

int help_gpu_renew(help_gpu* h, const double* d, help_renew_stats* rs) {
    if (!h || !d) return HELP_ERR_ARG;
    if (!h->device) { h->err = "renew: handle has no device state"; return HELP_ERR_STATE; }
    try {
        hipStream_t st = h->st;
        const int32_t m = h->m, n = h->n;
        HIPCHK(hipMemcpyAsync(h->d_d.p, d, sizeof(double) * n, hipMemcpyDefault, st));
        HIPCHK(hipEventRecord(h->ev[0], st));
        launch_dinv(h->d_d.p, h->d_dinv.p, n, st);
        HIPCHK(hipMemsetAsync(h->d_err.p, 0, sizeof(int32_t), st));
        if (h->use_pcg) {
            HIPCHK(hipEventRecord(h->ev[1], st));
            {
                ProfScope ps(h, HELP_K_PCG_BUILD, help::pcg_build_bytes(h));
                help::pcg_renew(h);
            }
            HIPCHK(hipEventRecord(h->ev[2], st));
            HIPCHK(hipStreamSynchronize(st));
            if (h->prof) collect_prof(h);
            h->factored = true;
            if (rs) {
                std::memset(rs, 0, sizeof(*rs));
                rs->ms_assemble = elapsed_ms(h->ev[0], h->ev[1]);
                rs->ms_precond = elapsed_ms(h->ev[1], h->ev[2]);
            }
            return HELP_OK;
        }
        // refinement has grown more expensive than the FP32 factor saves:
        // factor in FP64 from now on
        if (h->use_fp32 && !h->fp32_off && h->opts.refine_switch > 0 && h->cyc_solves > 0 &&
            (double)h->cyc_steps > h->opts.refine_switch * (double)h->cyc_solves)
            h->fp32_off = true;
        h->cyc_solves = h->cyc_steps = 0;
        const bool f32 = h->use_fp32 && !h->fp32_off;
        if (!f32) ensure_fp64_factor(h);
        assemble(h, f32);
        HIPCHK(hipEventRecord(h->ev[1], st));
        factor(h, f32);
        HIPCHK(hipEventRecord(h->ev[2], st));
        int32_t nerr = 0;
        HIPCHK(hipMemcpyAsync(&nerr, h->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (h->prof) collect_prof(h);
        h->last_tiny = nerr;
        h->factored = true;
        if (rs) {
            std::memset(rs, 0, sizeof(*rs));
            rs->ms_assemble = elapsed_ms(h->ev[0], h->ev[1]);
            rs->ms_factor = elapsed_ms(h->ev[1], h->ev[2]);
            rs->n_tiny_pivots = nerr;
            rs->factor_bits = f32 ? 32 : 64;
        }
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_solve(help_gpu* h, const double* d, const double* b1, const double* b2, double* x1,
                   double* x2, help_solve_stats* ss) {
    if (!h || !d || !b1 || !b2 || !x1 || !x2) return HELP_ERR_ARG;
    if (!h->device || !h->factored) { h->err = "solve: call renew first"; return HELP_ERR_STATE; }
    try {
        hipStream_t st = h->st;
        const int32_t m = h->m, n = h->n;
        help_solve_stats loc{};
        bool noconv = false;
        HIPCHK(hipMemcpyAsync(h->d_d.p, d, sizeof(double) * n, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(h->d_b1.p, b1, sizeof(double) * n, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(h->d_b2.p, b2, sizeof(double) * m, hipMemcpyDefault, st));
        HIPCHK(hipEventRecord(h->ev[0], st));
        // r = A (b1 ./ d) - b2, permuted (direct modes) or in the original order (PCG)
        {
            ProfScope ps(h, HELP_K_RHS, h->work_rhs);
            if (h->use_pcg) {
                launch_vdiv(n, h->d_b1.p, h->d_d.p, h->d_tn.p, st);
                launch_csr_split(h->rows, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_tn.p,
                                 h->d_rp.p, false, st);
                launch_sub(m, h->d_rp.p, h->d_b2.p, h->d_rp.p, st);
            } else {
                launch_rhs(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_b1.p, h->d_d.p, h->d_b2.p,
                           h->d_iperm.p, h->d_tn.p, h->d_rp.p, st);
            }
        }
        HIPCHK(hipEventRecord(h->ev[1], st));
        if (h->use_pcg) {
            // matrix-free PCG on A D^-1 A^T (UsePCG_CPU.cpp:91-178), original row order
            double rel = 0;
            int it;
            const double* x0 = nullptr;
            if (h->learned) {
                // the network's guess (PCG then runs to the same FP64 target)
                bool ok = h->net != nullptr && !(h->opts.fault_flags & HELP_FAULT_NET_NONFINITE);
                if (ok) {
                    help::mpnn_predict_device(h->net, h, h->d_d.p, h->d_rp.p, h->d_q.p);
                    HIPCHK(hipMemsetAsync(h->d_err.p, 0, sizeof(int32_t), st));
                    help::nn_check_finite(m, h->d_q.p, h->d_err.p, st);
                    int32_t bad = 0;
                    HIPCHK(hipMemcpyAsync(&bad, h->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
                    HIPCHK(hipStreamSynchronize(st));
                    ok = bad == 0;
                }
                if (ok) x0 = h->d_q.p;
                else loc.fallback = 1;
            }
            {
                ProfScope ps(h, HELP_K_PCG_SOLVE, 0.0);
                it = help::pcg_solve(h, h->d_rp.p, x0, h->d_x2.p, rel);
                ps.set_work(help::pcg_iteration_bytes(h) * (it + 1));
            }
            loc.pcg_iters = it;
            loc.rel_residual = rel;
            noconv = !(rel <= h->opts.pcg_tol);
        } else {
            // the factor's solution, refined by FP64 conjugate gradients on
            // (A D^-1 A^T + beta I) preconditioned with the same factor; the
            // FP64 mode takes the plain factor solve unless refine_fp64 > 0
            double rel = 0;
            int it = h->factor_fp32 ? pcg_factor<float>(h, rel, h->opts.refine_max)
                                    : pcg_factor<double>(h, rel, h->fp32_off ? h->opts.refine_max
                                                                             : h->opts.refine_fp64);
            if (h->factor_fp32) { h->cyc_solves += 1; h->cyc_steps += it; }
            if (h->factor_fp32 &&
                (!(rel <= h->opts.refine_tol) ||
                 (h->opts.fault_flags & (HELP_FAULT_FP32_BREAKDOWN | HELP_FAULT_REFINE_NOCONV)))) {
                // the FP32 factor no longer preconditions well enough: factor
                // in FP64 with the same d and keep FP64 for the rest of the run
                h->fp32_off = true;
                ensure_fp64_factor(h);
                assemble(h, false);
                factor(h, false);
                launch_scatter_perm(m, h->d_c.p, h->d_iperm.p, h->d_rp.p, st);   // permuted rhs again
                it += pcg_factor<double>(h, rel, h->opts.refine_max);
                loc.fallback = 1;
            }
            loc.refine_steps = it;
            loc.rel_residual = rel;
        }
        HIPCHK(hipEventRecord(h->ev[2], st));
        {
            ProfScope ps(h, HELP_K_BACKSUB, h->work_backsub);
            launch_backsub(n, h->d_colptr.p, h->d_rowind.p, h->d_val.p, h->d_b1.p, h->d_d.p, h->d_x2.p,
                           h->d_x1.p, st);
        }
        HIPCHK(hipEventRecord(h->ev[3], st));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(x1, h->d_x1.p, sizeof(double) * n, hipMemcpyDefault, st));
        HIPCHK(hipMemcpyAsync(x2, h->d_x2.p, sizeof(double) * m, hipMemcpyDefault, st));
        HIPCHK(hipStreamSynchronize(st));
        if (h->rel_deferred) {
            const double rr0 = h->h_pin[0], rr = h->h_pin[1];
            loc.rel_residual = rr0 == 0.0 ? 0.0 : std::sqrt(rr / rr0);
            h->rel_deferred = false;
        }
        if (h->prof) collect_prof(h);
        loc.ms_rhs = elapsed_ms(h->ev[0], h->ev[1]);
        loc.ms_solve = elapsed_ms(h->ev[1], h->ev[2]);
        loc.ms_back = elapsed_ms(h->ev[2], h->ev[3]);
        if (ss) *ss = loc;
        if (noconv) {
            h->err = "solve: PCG stopped at pcg_max_iter = " + std::to_string(h->opts.pcg_max_iter) +
                     " with relative residual " + std::to_string(loc.rel_residual) + " > pcg_tol";
            return HELP_ERR_NOCONV;
        }
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_solve2(help_gpu* h, const double* d, const double* b1a, const double* b2a, double* x1a,
                    double* x2a, const double* b1b, const double* b2b, double* x1b, double* x2b,
                    help_solve_stats* ssa, help_solve_stats* ssb) {
    if (!h || !d || !b1a || !b2a || !x1a || !x2a || !b1b || !b2b || !x1b || !x2b) return HELP_ERR_ARG;
    if (!h->device || !h->factored) { h->err = "solve2: call renew first"; return HELP_ERR_STATE; }
    const int max_it = h->fp32_off ? h->opts.refine_max : h->opts.refine_fp64;
    const bool lockstep = !h->use_pcg && h->factor_fp32 && h->opts.refine_max > 0 &&
                          !(h->opts.fault_flags & (HELP_FAULT_FP32_BREAKDOWN | HELP_FAULT_REFINE_NOCONV));
    if (lockstep) {
        // FP32 factor: both refinements in lockstep over shared triangular
        // solves (pcg_factor2).  When both reach refine_tol the results are
        // those of two single solves bit for bit; otherwise the two single
        // solves run after all (the first failure switches the run to the
        // FP64 factor), so a failure behaves exactly as before.
        try {
            hipStream_t st = h->st;
            const int32_t m = h->m, n = h->n;
            const double* b1[2] = {b1a, b1b};
            const double* b2[2] = {b2a, b2b};
            double* x1[2] = {x1a, x1b};
            double* x2[2] = {x2a, x2b};
            HIPCHK(hipMemcpyAsync(h->d_d.p, d, sizeof(double) * n, hipMemcpyDefault, st));
            for (int r = 0; r < 2; ++r) {
                HIPCHK(hipMemcpyAsync(h->d_b1.p + (size_t)r * n, b1[r], sizeof(double) * n, hipMemcpyDefault, st));
                HIPCHK(hipMemcpyAsync(h->d_b2.p + (size_t)r * m, b2[r], sizeof(double) * m, hipMemcpyDefault, st));
            }
            HIPCHK(hipEventRecord(h->ev[0], st));
            {
                ProfScope ps(h, HELP_K_RHS, 2 * h->work_rhs);
                for (int r = 0; r < 2; ++r)
                    launch_rhs(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_b1.p + (size_t)r * n,
                               h->d_d.p, h->d_b2.p + (size_t)r * m, h->d_iperm.p, h->d_tn.p,
                               h->d_rp.p + (size_t)r * m, st);
            }
            HIPCHK(hipEventRecord(h->ev[1], st));
            double rel[2];
            int it[2];
            pcg_factor2<float>(h, rel, it, h->opts.refine_max);
            if (rel[0] <= h->opts.refine_tol && rel[1] <= h->opts.refine_tol) {
                h->cyc_solves += 2;
                h->cyc_steps += it[0] + it[1];
                HIPCHK(hipEventRecord(h->ev[2], st));
                {
                    ProfScope ps(h, HELP_K_BACKSUB, 2 * h->work_backsub);
                    for (int r = 0; r < 2; ++r)
                        launch_backsub(n, h->d_colptr.p, h->d_rowind.p, h->d_val.p, h->d_b1.p + (size_t)r * n,
                                       h->d_d.p, h->d_x2.p + (size_t)r * m, h->d_x1.p + (size_t)r * n, st);
                }
                HIPCHK(hipEventRecord(h->ev[3], st));
                HIPCHK(hipGetLastError());
                for (int r = 0; r < 2; ++r) {
                    HIPCHK(hipMemcpyAsync(x1[r], h->d_x1.p + (size_t)r * n, sizeof(double) * n, hipMemcpyDefault, st));
                    HIPCHK(hipMemcpyAsync(x2[r], h->d_x2.p + (size_t)r * m, sizeof(double) * m, hipMemcpyDefault, st));
                }
                HIPCHK(hipStreamSynchronize(st));
                if (h->prof) collect_prof(h);
                help_solve_stats loc{};
                loc.ms_rhs = elapsed_ms(h->ev[0], h->ev[1]);
                loc.ms_solve = elapsed_ms(h->ev[1], h->ev[2]);
                loc.ms_back = elapsed_ms(h->ev[2], h->ev[3]);
                help_solve_stats* ss[2] = {ssa, ssb};
                for (int r = 0; r < 2; ++r) {
                    if (!ss[r]) continue;
                    *ss[r] = loc;
                    ss[r]->refine_steps = it[r];
                    ss[r]->rel_residual = rel[r];
                }
                return HELP_OK;
            }
        } catch (const std::exception& e) {
            return fail(h, e);
        }
    }
    if (h->use_pcg || h->factor_fp32 || max_it > 0) {
        // iterative paths decide per right-hand side: two single solves
        // (a PCG solve stopped at its cap still leaves an iterate: the
        // second system is solved too and the first failure is returned)
        const int sa = help_gpu_solve(h, d, b1a, b2a, x1a, x2a, ssa);
        if (sa != HELP_OK && sa != HELP_ERR_NOCONV) return sa;
        const int sb = help_gpu_solve(h, d, b1b, b2b, x1b, x2b, ssb);
        return sa != HELP_OK ? sa : sb;
    }
    try {
        hipStream_t st = h->st;
        const int32_t m = h->m, n = h->n;
        const double* b1[2] = {b1a, b1b};
        const double* b2[2] = {b2a, b2b};
        double* x1[2] = {x1a, x1b};
        double* x2[2] = {x2a, x2b};
        HIPCHK(hipMemcpyAsync(h->d_d.p, d, sizeof(double) * n, hipMemcpyDefault, st));
        for (int r = 0; r < 2; ++r) {
            HIPCHK(hipMemcpyAsync(h->d_b1.p + (size_t)r * n, b1[r], sizeof(double) * n, hipMemcpyDefault, st));
            HIPCHK(hipMemcpyAsync(h->d_b2.p + (size_t)r * m, b2[r], sizeof(double) * m, hipMemcpyDefault, st));
        }
        HIPCHK(hipEventRecord(h->ev[0], st));
        {
            ProfScope ps(h, HELP_K_RHS, 2 * h->work_rhs);
            for (int r = 0; r < 2; ++r)
                launch_rhs(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_b1.p + (size_t)r * n, h->d_d.p,
                           h->d_b2.p + (size_t)r * m, h->d_iperm.p, h->d_tn.p, h->d_rp.p + (size_t)r * m, st);
        }
        HIPCHK(hipEventRecord(h->ev[1], st));
        // the FP64 factor solve of both right-hand sides, then each one's
        // residual norms (reported only, as in the single solve)
        for (int r = 0; r < 2; ++r)
            launch_gather_perm(m, h->d_rp.p + (size_t)r * m, h->d_iperm.p, h->d_c.p + (size_t)r * m, st);
        run_trisolve<double>(h, h->d_L64.p, h->d_Dinv64.p, h->d_rp.p, 2);
        launch_dinv(h->d_d.p, h->d_dinv.p, n, st);
        for (int r = 0; r < 2; ++r) {
            double* x = h->d_x2.p + (size_t)r * m;
            const double* rhs = h->d_c.p + (size_t)r * m;
            launch_gather_perm(m, h->d_rp.p + (size_t)r * m, h->d_iperm.p, x, st);
            launch_normal_matvec(m, n, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_colptr.p,
                                 h->d_rowind.p, h->d_val.p, h->d_dinv.p, h->opts.beta, x, h->d_tn.p, h->d_q.p,
                                 st);
            launch_sub(m, rhs, h->d_q.p, h->d_r.p, st);
            launch_dot(rhs, rhs, m, h->d_part.p, h->d_scal.p + 2 + 2 * r, st);
            launch_dot(h->d_r.p, h->d_r.p, m, h->d_part.p, h->d_scal.p + 3 + 2 * r, st);
        }
        HIPCHK(hipMemcpyAsync(h->h_pin, h->d_scal.p + 2, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(h->ev[2], st));
        {
            ProfScope ps(h, HELP_K_BACKSUB, 2 * h->work_backsub);
            for (int r = 0; r < 2; ++r)
                launch_backsub(n, h->d_colptr.p, h->d_rowind.p, h->d_val.p, h->d_b1.p + (size_t)r * n, h->d_d.p,
                               h->d_x2.p + (size_t)r * m, h->d_x1.p + (size_t)r * n, st);
        }
        HIPCHK(hipEventRecord(h->ev[3], st));
        HIPCHK(hipGetLastError());
        for (int r = 0; r < 2; ++r) {
            HIPCHK(hipMemcpyAsync(x1[r], h->d_x1.p + (size_t)r * n, sizeof(double) * n, hipMemcpyDefault, st));
            HIPCHK(hipMemcpyAsync(x2[r], h->d_x2.p + (size_t)r * m, sizeof(double) * m, hipMemcpyDefault, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        if (h->prof) collect_prof(h);
        help_solve_stats* ss[2] = {ssa, ssb};
        for (int r = 0; r < 2; ++r) {
            if (!ss[r]) continue;
            help_solve_stats loc{};
            const double rr0 = h->h_pin[2 * r], rr = h->h_pin[2 * r + 1];
            loc.rel_residual = rr0 == 0.0 ? 0.0 : std::sqrt(rr / rr0);
            // the shared launches' times, split evenly
            loc.ms_rhs = 0.5 * elapsed_ms(h->ev[0], h->ev[1]);
            loc.ms_solve = 0.5 * elapsed_ms(h->ev[1], h->ev[2]);
            loc.ms_back = 0.5 * elapsed_ms(h->ev[2], h->ev[3]);
            *ss[r] = loc;
        }
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_symbolic_info(const help_gpu* h, help_symbolic_info* info) {
    if (!h || !info) return HELP_ERR_ARG;
    if (h->use_pcg) return HELP_ERR_STATE;   // no ordering or factor in PCG mode
    const Supernodal& S = h->S;
    std::memset(info, 0, sizeof(*info));
    info->nnz_adat_upper = (int64_t)h->P.colind.size();
    info->nnz_l = S.loff[S.nsuper];
    info->nnz_l_exact = S.nnz_l_exact;
    info->n_supernodes = S.nsuper;
    info->n_levels = S.nlevels;
    int32_t mx = 0;
    for (int32_t s = 0; s < S.nsuper; ++s) mx = std::max(mx, S.snrow[s]);
    info->max_front = mx;
    info->factor_flops = S.flops;
    info->n_pairs = (int64_t)h->P.pair_p.size();
    return HELP_OK;
}

int help_gpu_symbolic_arrays(const help_gpu* h, int32_t* perm, int32_t* parent, int32_t* colcount) {
    if (!h) return HELP_ERR_ARG;
    if (h->use_pcg) return HELP_ERR_STATE;   // no ordering or factor in PCG mode
    if (perm) std::memcpy(perm, h->S.perm.data(), sizeof(int32_t) * h->m);
    if (parent) std::memcpy(parent, h->S.parent.data(), sizeof(int32_t) * h->m);
    if (colcount) std::memcpy(colcount, h->S.colcount.data(), sizeof(int32_t) * h->m);
    return HELP_OK;
}

int help_gpu_adat_pattern(const help_gpu* h, int64_t* rowptr, int32_t* colind) {
    if (!h) return HELP_ERR_ARG;
    if (h->use_pcg) return HELP_ERR_STATE;   // no ordering or factor in PCG mode
    if (rowptr) std::memcpy(rowptr, h->P.rowptr.data(), sizeof(int64_t) * (h->m + 1));
    if (colind) std::memcpy(colind, h->P.colind.data(), sizeof(int32_t) * h->P.colind.size());
    return HELP_OK;
}

int help_gpu_adat_values(help_gpu* h, const double* d, double* vals) {
    if (!h || !d || !vals) return HELP_ERR_ARG;
    if (!h->device || h->use_pcg) { h->err = "adat_values: no device state or PCG mode"; return HELP_ERR_STATE; }
    try {
        hipStream_t st = h->st;
        int64_t nslot = (int64_t)h->P.colind.size();
        h->d_slotvals.alloc(std::max<int64_t>(1, nslot));
        HIPCHK(hipMemcpyAsync(h->d_d.p, d, sizeof(double) * h->n, hipMemcpyHostToDevice, st));
        launch_dinv(h->d_d.p, h->d_dinv.p, h->n, st);
        launch_adat_assemble(h->d_pairptr.p, h->d_pair_p.p, h->d_pair_q.p, h->d_colof.p, h->d_val.p,
                             h->d_dinv.p, h->nnz, h->d_wnz.p, nslot, nullptr, h->d_slotvals.p, nullptr, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(vals, h->d_slotvals.p, sizeof(double) * nslot, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        h->d_slotvals.free();
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_set_warmstart(help_gpu* h, help_mpnn* net) {
    if (!h) return HELP_ERR_ARG;
    if (!h->learned) { h->err = "set_warmstart: handle is not in the learned PCG mode"; return HELP_ERR_STATE; }
    h->net = net;
    return HELP_OK;
}

int help_gpu_pcg_factor(help_gpu* h, int32_t* K, int32_t* perm, double* D, double* S, double* L) {
    if (!h) return HELP_ERR_ARG;
    if (!h->use_pcg || !h->device || !h->factored) { h->err = "pcg_factor: no PCG preconditioner"; return HELP_ERR_STATE; }
    try {
        help::pcg_download(h, K, perm, D, S, L);
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_pcg_apply(help_gpu* h, const double* r, double* z) {
    if (!h || !r || !z) return HELP_ERR_ARG;
    if (!h->device || !h->use_pcg || !h->factored) { h->err = "pcg_apply: PCG mode after renew only"; return HELP_ERR_STATE; }
    try {
        hipStream_t st = h->st;
        const int32_t m = h->m;
        HIPCHK(hipMemcpyAsync(h->d_c.p, r, sizeof(double) * m, hipMemcpyDefault, st));
        help::pcg_precond(h, h->d_c.p, h->d_z.p);
        HIPCHK(hipMemcpyAsync(z, h->d_z.p, sizeof(double) * m, hipMemcpyDefault, st));
        HIPCHK(hipStreamSynchronize(st));
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

static bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();   // a host pointer leaves an error behind: clear it
    return e == hipSuccess && a.type == hipMemoryTypeDevice;
}

int help_gpu_pcg_apply_noalloc(help_gpu* h, const double* r, double* z) {
    if (!h || !r || !z) return HELP_ERR_ARG;
    if (!h->device || !h->use_pcg || !h->factored) { h->err = "pcg_apply_noalloc: PCG mode after renew only"; return HELP_ERR_STATE; }
    if (!is_device_ptr(r) || !is_device_ptr(z)) {
        h->err = "pcg_apply_noalloc: r and z must be device memory";
        return HELP_ERR_ARG;
    }
    try {
        help::pcg_precond_noalloc(h, r, z);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->st));
        return HELP_OK;
    } catch (const HipError& e) {
        return fail(h, e);
    } catch (const std::exception& e) {
        h->err = e.what();
        return HELP_ERR_ARG;
    }
}

int help_gpu_spmv(help_gpu* h, int32_t trans, const double* x, double* y) {
    if (!h || !x || !y) return HELP_ERR_ARG;
    if (!h->device) { h->err = "spmv: no device state"; return HELP_ERR_STATE; }
    try {
        hipStream_t st = h->st;
        const int32_t m = h->m, n = h->n;
        if (!trans) {
            HIPCHK(hipMemcpyAsync(h->d_b1.p, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
            launch_spmv_csr(m, h->d_rowptr.p, h->d_colind.p, h->d_csrval.p, h->d_b1.p, h->d_b2.p, st);
            HIPCHK(hipMemcpyAsync(y, h->d_b2.p, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        } else {
            HIPCHK(hipMemcpyAsync(h->d_b2.p, x, sizeof(double) * m, hipMemcpyHostToDevice, st));
            launch_spmv_csc_t(n, h->d_colptr.p, h->d_rowind.p, h->d_val.p, h->d_b2.p, h->d_b1.p, st);
            HIPCHK(hipMemcpyAsync(y, h->d_b1.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_factor_column(help_gpu* h, int32_t j, int32_t* rows, double* vals, int32_t* nr) {
    if (!h || !rows || !vals || !nr || j < 0 || j >= h->m) return HELP_ERR_ARG;
    if (!h->device || !h->factored || h->use_pcg) { h->err = "factor_column: no factor"; return HELP_ERR_STATE; }
    try {
        const Supernodal& S = h->S;
        int32_t s = S.col2sn[j];
        int32_t c = S.snrow[s], jj = j - S.sfirst[s];
        int64_t off = S.loff[s] + (int64_t)jj * c;
        int32_t cnt = c - jj;
        if (h->factor_fp32) {
            std::vector<float> tmp(cnt);
            HIPCHK(hipMemcpy(tmp.data(), h->d_L32.p + off + jj, sizeof(float) * cnt, hipMemcpyDeviceToHost));
            for (int32_t t = 0; t < cnt; ++t) vals[t] = tmp[t];
        } else {
            HIPCHK(hipMemcpy(vals, h->d_L64.p + off + jj, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        }
        for (int32_t t = 0; t < cnt; ++t) rows[t] = S.rowidx[S.rowoff[s] + jj + t];
        *nr = cnt;
        return HELP_OK;
    } catch (const std::exception& e) {
        return fail(h, e);
    }
}

int help_gpu_profile(help_gpu* h, int32_t enable) {
    if (!h) return HELP_ERR_ARG;
    h->prof = enable != 0;
    for (auto& k : h->kstat) k = help_gpu::KStat();
    for (auto& k : h->kstat32) k = help_gpu::KStat();
    h->pending.clear();
    return HELP_OK;
}

int help_gpu_kernel_stats(const help_gpu* h, int32_t kind, double* ms, int64_t* launches,
                          double* work, const char** name) {
    if (!h || kind < 0 || kind >= HELP_K_COUNT) return HELP_ERR_ARG;
    if (ms) *ms = h->kstat[kind].ms;
    if (launches) *launches = h->kstat[kind].launches;
    if (work) *work = h->kstat[kind].work;
    if (name) *name = kKindName[kind];
    return HELP_OK;
}

int help_device_synchronize(void) {
    return hipDeviceSynchronize() == hipSuccess ? HELP_OK : HELP_ERR_HIP;
}

int help_gpu_supernodes(const help_gpu* h, int32_t* sfirst, int32_t* snrow, int32_t* level,
                        int32_t* tiled) {
    if (!h) return HELP_ERR_ARG;
    if (h->use_pcg) return HELP_ERR_STATE;
    const Supernodal& S = h->S;
    if (sfirst) std::memcpy(sfirst, S.sfirst.data(), sizeof(int32_t) * (S.nsuper + 1));
    if (snrow) std::memcpy(snrow, S.snrow.data(), sizeof(int32_t) * S.nsuper);
    if (level) std::memcpy(level, S.level.data(), sizeof(int32_t) * S.nsuper);
    if (tiled) {
        for (int32_t s = 0; s < S.nsuper; ++s) tiled[s] = 1;
        for (int32_t s : h->list_small) tiled[s] = 0;
    }
    return HELP_OK;
}

int help_selftest_mfma(int32_t fp64, const double* a, const double* b, double* out) {
    if (!a || !b || !out) return HELP_ERR_ARG;
    try {
        if (help_device_count() == 0) throw HipError("no HIP device visible", HELP_ERR_NOGPU);
        if (fp64) {
            DBuf<double> da, db, dout;
            da.alloc(64); db.alloc(64); dout.alloc(256);
            HIPCHK(hipMemcpy(da.p, a, 64 * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(db.p, b, 64 * sizeof(double), hipMemcpyHostToDevice));
            launch_mfma_probe<double>(da.p, db.p, dout.p, nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpy(out, dout.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> fa(a, a + 64), fb(b, b + 64), fo(256);
            DBuf<float> da, db, dout;
            da.alloc(64); db.alloc(64); dout.alloc(256);
            HIPCHK(hipMemcpy(da.p, fa.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(db.p, fb.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
            launch_mfma_probe<float>(da.p, db.p, dout.p, nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpy(fo.data(), dout.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
            for (int i = 0; i < 256; ++i) out[i] = fo[i];
        }
        return HELP_OK;
    } catch (const HipError& e) {
        g_create_error = e.what();
        return e.code;
    }
}

// ------------------------------------------------------------------ backend
struct GpuBackendCtx {
    help_opts opts;
    help_gpu* h = nullptr;
    int32_t last_status = 0;
    help_gpu::KStat kstat[HELP_K_COUNT], kstat32[HELP_K_COUNT];
    int64_t n_solve = 0, refine_steps = 0;
    double max_rel_residual = 0;
    int64_t pcg_iters = 0;
    help_mpnn* net = nullptr;   // warm start for the learned PCG mode (not owned)
};

static void gb_release(GpuBackendCtx* c) {
    if (!c->h) return;
    for (int k = 0; k < HELP_K_COUNT; ++k) {
        c->kstat[k].ms += c->h->kstat[k].ms;
        c->kstat[k].work += c->h->kstat[k].work;
        c->kstat[k].launches += c->h->kstat[k].launches;
        c->kstat32[k].ms += c->h->kstat32[k].ms;
        c->kstat32[k].work += c->h->kstat32[k].work;
        c->kstat32[k].launches += c->h->kstat32[k].launches;
    }
    help_gpu_destroy(c->h);
    c->h = nullptr;
}

static int gb_construct(void* ctx, const help_lp* lp) {
    auto* c = (GpuBackendCtx*)ctx;
    gb_release(c);
    int st = help_gpu_create(lp->m, lp->n, lp->nnz, lp->colptr, lp->rowind, lp->val, &c->opts, &c->h);
    if (st == HELP_OK && c->opts.mode == HELP_MODE_PCG_LEARNED) st = help_gpu_set_warmstart(c->h, c->net);
    return st;
}
static int gb_destruct(void* ctx) {
    gb_release((GpuBackendCtx*)ctx);
    return 0;
}
static int gb_renew(void* ctx, const double* d) {
    auto* c = (GpuBackendCtx*)ctx;
    return help_gpu_renew(c->h, d, nullptr);
}
static int gb_solve(void* ctx, const double* d, const double* b1, const double* b2, double* x1,
                    double* x2) {
    auto* c = (GpuBackendCtx*)ctx;
    help_solve_stats ss{};
    int st = help_gpu_solve(c->h, d, b1, b2, x1, x2, &ss);
    if (st == HELP_OK || st == HELP_ERR_NOCONV) {
        c->n_solve += 1;
        c->refine_steps += ss.refine_steps;
        c->pcg_iters += ss.pcg_iters;
        c->max_rel_residual = std::max(c->max_rel_residual, ss.rel_residual);
    }
    return st;
}

int help_gpu_backend_solve2(help_backend* be, const double* d, const double* b1a, const double* b2a,
                            double* x1a, double* x2a, const double* b1b, const double* b2b, double* x1b,
                            double* x2b) {
    if (!be || !be->ctx) return HELP_ERR_ARG;
    auto* c = (GpuBackendCtx*)be->ctx;
    help_solve_stats ss[2]{};
    int st = help_gpu_solve2(c->h, d, b1a, b2a, x1a, x2a, b1b, b2b, x1b, x2b, &ss[0], &ss[1]);
    if (st == HELP_OK || st == HELP_ERR_NOCONV) {
        for (const help_solve_stats& s : ss) {
            c->n_solve += 1;
            c->refine_steps += s.refine_steps;
            c->pcg_iters += s.pcg_iters;
            c->max_rel_residual = std::max(c->max_rel_residual, s.rel_residual);
        }
    }
    return st;
}

int help_gpu_backend_new(const help_opts* opts, help_backend* out) {
    if (!out) return HELP_ERR_ARG;
    auto* c = new (std::nothrow) GpuBackendCtx();
    if (!c) return HELP_ERR_NOMEM;
    if (opts) c->opts = *opts; else help_opts_default(&c->opts);
    c->opts.symbolic_only = 0;
    out->ctx = c;
    out->construct = gb_construct;
    out->destruct = gb_destruct;
    out->renew = gb_renew;
    out->solve = gb_solve;
    return HELP_OK;
}

void help_gpu_backend_free(help_backend* be) {
    if (!be || !be->ctx) return;
    auto* c = (GpuBackendCtx*)be->ctx;
    help_gpu_destroy(c->h);
    delete c;
    be->ctx = nullptr;
}

help_gpu* help_gpu_backend_handle(const help_backend* be) {
    if (!be || !be->ctx) return nullptr;
    return ((GpuBackendCtx*)be->ctx)->h;
}

int help_gpu_backend_set_warmstart(help_backend* be, help_mpnn* net) {
    if (!be || !be->ctx) return HELP_ERR_ARG;
    ((GpuBackendCtx*)be->ctx)->net = net;
    return HELP_OK;
}

int64_t help_gpu_backend_pcg_iterations(const help_backend* be) {
    if (!be || !be->ctx) return -1;
    return ((GpuBackendCtx*)be->ctx)->pcg_iters;
}

int help_gpu_backend_solve_stats(const help_backend* be, int64_t* n_solve, int64_t* refine_steps,
                                 double* max_rel_residual) {
    if (!be || !be->ctx) return HELP_ERR_ARG;
    auto* c = (GpuBackendCtx*)be->ctx;
    if (n_solve) *n_solve = c->n_solve;
    if (refine_steps) *refine_steps = c->refine_steps;
    if (max_rel_residual) *max_rel_residual = c->max_rel_residual;
    return HELP_OK;
}

int help_gpu_backend_kernel_stats(const help_backend* be, int32_t kind, double* ms,
                                  int64_t* launches, double* work) {
    return help_gpu_backend_kernel_stats_prec(be, kind, 0, ms, launches, work);
}

int help_gpu_backend_kernel_stats_prec(const help_backend* be, int32_t kind, int32_t fp32_only, double* ms,
                                       int64_t* launches, double* work) {
    if (!be || !be->ctx || kind < 0 || kind >= HELP_K_COUNT) return HELP_ERR_ARG;
    auto* c = (GpuBackendCtx*)be->ctx;
    const help_gpu::KStat* own = fp32_only ? c->kstat32 : c->kstat;
    double t = own[kind].ms, w = own[kind].work;
    int64_t n = own[kind].launches;
    if (c->h) {
        const help_gpu::KStat* live = fp32_only ? c->h->kstat32 : c->h->kstat;
        t += live[kind].ms; w += live[kind].work; n += live[kind].launches;
    }
    if (ms) *ms = t;
    if (launches) *launches = n;
    if (work) *work = w;
    return HELP_OK;
}

}  // extern "C"
