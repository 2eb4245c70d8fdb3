// Single large circuit instance (tens of thousands of unknowns): the host symbolic phase of its sparse LU
// and the CPU reference of the numeric phase.  The batched engine keeps one instance's whole matrix in a
// workgroup's LDS; a matrix of this size lives in HBM and is factored by the multifrontal method over a
// fill-reducing order instead (sparse21/mod.rs:647-991 is the factorisation it replaces).
#pragma once
#include <cstdint>
#include <vector>

namespace s21 {

// Static pivoting, ordering and supernodal structure of one sparse matrix pattern.
//
// Index spaces: external (the circuit's variables, rows and columns of A) and internal k = 0..n-1 (the
// elimination order).  Step k eliminates column perm[k] with pivot row prow[k]; internal entry (a, b) is
// A(prow[a], perm[b]).  Fronts: supernode s owns internal columns [sn_first[s], sn_first[s+1]) and the
// row set R_s (internal indices above them, ascending); its front is the dense F x F matrix over the
// index list I_s = [own columns, R_s], stored column-major at front_off[s] in one pool.
struct LargeSym {
    int n = 0;
    std::vector<int> perm, prow;    // [n] internal step -> external column / external pivot row
    std::vector<int> icol, irow;    // [n] external column / row -> internal index
    int ns = 0;
    std::vector<int> sn_first;      // [ns + 1]
    std::vector<int> sn_parent;     // [ns], -1 for a root
    std::vector<int64_t> r_ptr;     // [ns + 1] into r_idx
    std::vector<int> r_idx;         // R_s, internal indices, ascending
    std::vector<int64_t> front_off; // [ns + 1] offsets (entries) into the front pool
    // children: sn_child_ptr / sn_child list each supernode's children in ascending order; ext_ptr /
    // ext_map give, per supernode, the positions of its R_s entries within its parent's index list
    std::vector<int> child_ptr, child;
    std::vector<int64_t> ext_ptr;
    std::vector<int> ext_map;
    // level schedule: supernodes of level l (leaves 0) are lvl_sn[lvl_ptr[l] .. lvl_ptr[l+1])
    std::vector<int> lvl_ptr, lvl_sn;
    // statistics
    int64_t nnz_a = 0, nnz_lu = 0, pool = 0;
    double flops = 0.0;
    int max_front = 0;
};

// Builds the structure of the n x n pattern (rows[e], cols[e]) (duplicates allowed).  `vals` (optional,
// one per entry) are the first Newton iteration's values: the static pivot rows are chosen among entries
// whose value is nonzero when possible.  relax: supernodes merge while the explicit zeros they add stay
// within this fraction of the merged front (0 = fundamental supernodes only).
void large_symbolic(int n, const std::vector<int>& rows, const std::vector<int>& cols, const std::vector<double>* vals,
                    LargeSym& S, double relax = 0.25);

// Offset (into the front pool) of internal entry (a, b), or -1 when the structure has no such entry.
int64_t large_entry_offset(const LargeSym& S, int a, int b);

// CPU reference of the numeric phase (FP64), same operations in the same order as the GPU kernels:
// assembles A's values into the fronts, factors, and solves A x = b.  Used by the tests.
void large_factor_host(const LargeSym& S, const std::vector<int>& rows, const std::vector<int>& cols,
                       const std::vector<double>& vals, std::vector<double>& pool);
void large_solve_host(const LargeSym& S, const std::vector<double>& pool, const std::vector<double>& b,
                      std::vector<double>& x);

// Approximate minimum degree order of the symmetric pattern adj (CSR without the diagonal): order[k] is
// the vertex eliminated k-th.
void amd_order(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& order);

}  // namespace s21
