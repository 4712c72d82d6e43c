// Host symbolic analysis of the normal matrix M = A D^-1 A^T + beta I.
//
// Runs once per LP (the reference keeps CHOLMOD's analysis in
// CHOL_Fac_Backup, UseCholmod.cpp:107-126): the upper pattern of A A^T in
// HELP's slot order (Helper.cpp:273-329), a fill-reducing approximate
// minimum degree ordering, the elimination tree, its postorder, exact column
// counts, relaxed supernodes and the index maps the GPU kernels consume.
#pragma once
#include <cstdint>
#include <vector>

namespace help {

struct CscMatrix {             // A, m x n, rows ascending within columns
    int32_t m = 0, n = 0;
    std::vector<int64_t> colptr;
    std::vector<int32_t> rowind;
    std::vector<double> val;
};

struct CsrMatrix {             // A by rows, columns ascending within rows
    int32_t m = 0, n = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> colind;
    std::vector<int64_t> cscpos;   // position of each entry in the CSC arrays
};

CsrMatrix csc_to_csr(const CscMatrix& A);

// Upper pattern of A A^T in HELP's order: row i lists the diagonal first and
// then new columns k > i in discovery order (for each (i,j) in row i by
// ascending j, rows of column j from i downwards).  Also the contribution
// lists that define each value: slot s = (i,k) sums over j ascending the
// products (A_ij * dinv_j) * A_kj; pair t holds the CSC positions of A_ij
// and A_kj.
struct AdatPattern {
    std::vector<int64_t> rowptr;      // m+1
    std::vector<int32_t> colind;      // nnz_upper
    std::vector<int64_t> pairptr;     // nnz_upper+1
    std::vector<int32_t> pair_p;      // CSC position of A_ij
    std::vector<int32_t> pair_q;      // CSC position of A_kj
};
AdatPattern adat_pattern(const CscMatrix& A, const CsrMatrix& R, bool with_pairs);

// Approximate minimum degree ordering of a symmetric graph given as full
// adjacency (CSR, no self loops).  Returns perm (new -> old).
std::vector<int32_t> amd_order(int32_t n, const std::vector<int64_t>& adjptr,
                               const std::vector<int32_t>& adj, double dense_alpha = 10.0);

// Nested dissection by level-structure bisection, subgraphs of at most
// `leaf` vertices ordered by amd_order.  Returns perm (new -> old).
std::vector<int32_t> nd_order(int32_t n, const std::vector<int64_t>& adjptr,
                              const std::vector<int32_t>& adj, int32_t leaf = 2048);

struct Supernodal {
    int32_t m = 0;
    std::vector<int32_t> perm, iperm;        // new -> old, old -> new
    std::vector<int32_t> parent;             // column etree (new numbering)
    std::vector<int32_t> colcount;           // exact |L(:,j)|, incl. diagonal
    // supernodes (ascending = postorder)
    int32_t nsuper = 0;
    std::vector<int32_t> sfirst;             // nsuper+1, first column of each
    std::vector<int32_t> snrow;              // rows in the structure (incl. own cols)
    std::vector<int32_t> sparent;            // parent supernode or -1
    std::vector<int64_t> rowoff;             // nsuper+1 offsets into rowidx
    std::vector<int32_t> rowidx;             // sorted row structure per supernode
    std::vector<int64_t> loff;               // nsuper+1 offsets into L values (nrow x ncol, col-major)
    std::vector<int64_t> uoff;               // nsuper+1 offsets into update matrices ((nrow-ncol)^2)
    std::vector<int64_t> reloff;             // nsuper+1 offsets into relmap
    std::vector<int32_t> relmap;             // per update row: index in parent's structure
    std::vector<int32_t> childptr, child;    // children lists (ascending)
    std::vector<int32_t> level;              // 0 for leaves
    int32_t nlevels = 0;
    std::vector<int32_t> levelptr, levelsn;  // supernodes grouped by level
    std::vector<int32_t> col2sn;             // column -> supernode
    int64_t nnz_l_exact = 0;
    double flops = 0;                        // sum colcount^2
};

// Full symmetric adjacency (no diagonal, rows sorted) of the pattern of
// A A^T, i.e. of ADA^T: rows i and k are adjacent when they share a column.
void adat_adjacency(const CscMatrix& A, const CsrMatrix& R, std::vector<int64_t>& adjptr,
                    std::vector<int32_t>& adj);

// The adjacency relabelled by perm (new -> old): row k holds iperm[] of the
// neighbours of perm[k], in their order there; half < 0 keeps only the
// neighbours numbered below k (what the elimination tree reads), half > 0
// only those above k (what the column counts read).
void permute_graph(int32_t m, const std::vector<int64_t>& adjptr, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& perm, const std::vector<int32_t>& iperm,
                   std::vector<int64_t>& pp, std::vector<int32_t>& pj, int half = 0);

// Elimination tree of P M P^T (Liu's algorithm with path compression), from
// the permuted adjacency (pp, pj).
std::vector<int32_t> etree(int32_t m, const std::vector<int64_t>& pp, const std::vector<int32_t>& pj);
std::vector<int32_t> postorder(const std::vector<int32_t>& parent);
// column counts from the permuted adjacency of a postordered tree
std::vector<int32_t> colcounts(int32_t m, const std::vector<int64_t>& pp, const std::vector<int32_t>& pj,
                               const std::vector<int32_t>& parent);
// the same counts by walking every row subtree (O(nnz(L))); test reference
std::vector<int32_t> colcounts_walk(int32_t m, const std::vector<int64_t>& adjptr,
                               const std::vector<int32_t>& adj, const std::vector<int32_t>& perm,
                               const std::vector<int32_t>& iperm, const std::vector<int32_t>& parent);

struct RelaxParams {          // merge child into parent when the merged supernode
    int32_t nrelax[3] = {4, 16, 48};       // has fewer columns than nrelax[k] and
    double zrelax[3] = {0.8, 0.1, 0.05};   // a zero fraction below zrelax[k]
    int32_t max_cols = 256;                // never merge beyond this many columns
    int32_t max_sn_cols = 2048;            // split any supernode wider than this
};

// Full analysis: ordering (perm_in may be empty -> AMD), etree, postorder,
// column counts, relaxed supernodes, structures, relative maps and levels.
Supernodal analyze(int32_t m, const std::vector<int64_t>& adjptr, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& perm_in, const RelaxParams& rp);

// Assembly maps of the numeric phase: factor position of every ADA^T slot,
// of every diagonal, the column of every nonzero of A, A's values in CSR
// order.  OpenMP-parallel.
void assembly_maps(const Supernodal& S, const AdatPattern& P, const CscMatrix& A, const CsrMatrix& R,
                   std::vector<int64_t>& slot_dest, std::vector<int64_t>& diagpos,
                   std::vector<int32_t>& colof, std::vector<double>& csrval);

}  // namespace help
