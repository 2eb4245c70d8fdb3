// Host symbolic phase of the single-large-instance path: static pivot rows (a maximum transversal), a
// fill-reducing order (approximate minimum degree on the row-matched pattern made symmetric), the
// elimination tree in postorder, supernodes and the multifrontal index maps; plus the CPU reference of
// the numeric phase the GPU kernels mirror.
//
// The reference searches Markowitz pivots at every factorisation (sparse21/mod.rs:647-863); restated,
// that search grows about as n^2.5 and takes over an hour at 45,000 unknowns.  This path instead fixes one
// order for the whole analysis, as the batched path does, but chooses it for fill, not by the reference's
// rules: Newton then converges to the same solution (the FP64 results agree with the reference's order to
// rounding), with a factorisation whose work is dominated by a few dense fronts.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_map>

#include "internal.h"
#include "large.h"

namespace s21 {

// ---------------------------------------------------------------------------------------------------
// Approximate minimum degree (Amestoy, Davis & Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996): quotient
// graph of variables and elements, element absorption, approximate external degrees from |Le \ Lp|,
// supervariables found by hashing the adjacency of the variables touched by each pivot.
// ---------------------------------------------------------------------------------------------------
void amd_order(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& order) {
    enum : uint8_t { VAR = 0, ELEM = 1, DEAD = 2, MERGED = 3 };
    std::vector<std::vector<int>> lst(n);  // variable: elen elements, then variables; element: variables
    std::vector<int> elen(n, 0), nv(n, 1), deg(n, 0), esize(n, 0), sv_parent(n, -1);
    std::vector<uint8_t> state(n, VAR);
    for (int i = 0; i < n; i++) {
        lst[i].assign(adj.begin() + ptr[i], adj.begin() + ptr[i + 1]);
        deg[i] = (int)lst[i].size();
    }
    // degree buckets (doubly linked)
    std::vector<int> head(n + 1, -1), nxt(n, -1), prv(n, -1);
    auto insert = [&](int i) {
        int d = std::min(std::max(deg[i], 0), n);
        nxt[i] = head[d];
        prv[i] = -1;
        if (head[d] >= 0) prv[head[d]] = i;
        head[d] = i;
    };
    auto remove = [&](int i) {
        int d = std::min(std::max(deg[i], 0), n);
        if (prv[i] >= 0)
            nxt[prv[i]] = nxt[i];
        else
            head[d] = nxt[i];
        if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
        nxt[i] = prv[i] = -1;
    };
    for (int i = 0; i < n; i++) insert(i);
    std::vector<int> flag(n, 0), wstamp(n, 0), w(n, 0);
    int mark = 0, wmark = 0, mindeg = 0;
    std::vector<int> pivots;  // principal pivots in elimination order
    int k = 0;
    std::vector<int> Lp;
    std::unordered_map<uint64_t, std::vector<int>> buckets;
    while (k < n) {
        while (mindeg <= n && head[mindeg] < 0) mindeg++;
        if (mindeg > n) break;
        const int p = head[mindeg];
        remove(p);
        pivots.push_back(p);
        k += nv[p];
        // ---- the new element: Lp = (A_p U union of Le over e in E_p) \ {p}; E_p absorbed ----
        ++mark;
        Lp.clear();
        int lpw = 0;
        auto take = [&](int v) {
            if (v != p && state[v] == VAR && nv[v] > 0 && flag[v] != mark) {
                flag[v] = mark;
                Lp.push_back(v);
                lpw += nv[v];
            }
        };
        for (int t = 0; t < elen[p]; t++) {
            const int e = lst[p][t];
            if (state[e] != ELEM) continue;
            for (int v : lst[e]) take(v);
            state[e] = DEAD;
            std::vector<int>().swap(lst[e]);
        }
        for (size_t t = elen[p]; t < lst[p].size(); t++) take(lst[p][t]);
        state[p] = ELEM;
        lst[p] = Lp;
        elen[p] = 0;
        esize[p] = lpw;
        for (int i : Lp) remove(i);
        // ---- |Le \ Lp| for the elements next to Lp ----
        ++wmark;
        for (int i : Lp)
            for (int t = 0; t < elen[i]; t++) {
                const int e = lst[i][t];
                if (state[e] != ELEM || e == p) continue;
                if (wstamp[e] != wmark) {
                    wstamp[e] = wmark;
                    w[e] = esize[e];
                }
                w[e] -= nv[i];
            }
        // ---- prune the lists of Lp and bound their degrees ----
        for (int i : Lp) {
            std::vector<int> ne{p}, nvar;
            int de = 0, da = 0;
            for (int t = 0; t < elen[i]; t++) {
                const int e = lst[i][t];
                if (state[e] != ELEM || e == p) continue;
                if (w[e] <= 0) {  // Le within Lp: absorbed into the new element
                    state[e] = DEAD;
                    std::vector<int>().swap(lst[e]);
                    continue;
                }
                ne.push_back(e);
                de += w[e];
            }
            for (size_t t = elen[i]; t < lst[i].size(); t++) {
                const int v = lst[i][t];
                if (v == p || state[v] != VAR || nv[v] == 0 || flag[v] == mark) continue;
                nvar.push_back(v);
                da += nv[v];
            }
            elen[i] = (int)ne.size();
            ne.insert(ne.end(), nvar.begin(), nvar.end());
            lst[i].swap(ne);
            const int ext = lpw - nv[i];
            int d = std::min({n - k - nv[i], deg[i] + ext, da + ext + de});
            deg[i] = std::max(d, 0);
        }
        // ---- supervariables: members of Lp with identical lists ----
        buckets.clear();
        for (int i : Lp) {
            uint64_t h = 0;
            for (int v : lst[i]) h += (uint64_t)v * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
            h ^= (uint64_t)lst[i].size() << 40;
            buckets[h].push_back(i);
        }
        for (auto& kv : buckets) {
            std::vector<int>& B = kv.second;
            if (B.size() < 2) continue;
            std::vector<std::vector<int>> sorted(B.size());
            for (size_t a = 0; a < B.size(); a++) {
                sorted[a] = lst[B[a]];
                std::sort(sorted[a].begin(), sorted[a].end());
            }
            for (size_t a = 0; a < B.size(); a++) {
                const int i = B[a];
                if (nv[i] == 0) continue;
                for (size_t b = a + 1; b < B.size(); b++) {
                    const int j = B[b];
                    if (nv[j] == 0 || elen[i] != elen[j] || sorted[a] != sorted[b]) continue;
                    // j joins i: ordered right after it
                    nv[i] += nv[j];
                    deg[i] -= nv[j];
                    nv[j] = 0;
                    state[j] = MERGED;
                    sv_parent[j] = i;
                    std::vector<int>().swap(lst[j]);
                }
            }
        }
        for (int i : Lp)
            if (nv[i] > 0) {
                deg[i] = std::max(std::min(deg[i], n - k - nv[i]), 0);
                insert(i);
                mindeg = std::min(mindeg, deg[i]);
            }
    }
    // members follow their principal variable
    std::vector<std::vector<int>> members(n);
    for (int j = 0; j < n; j++)
        if (sv_parent[j] >= 0) members[sv_parent[j]].push_back(j);
    order.clear();
    std::vector<int> stack;
    for (int p : pivots) {
        stack.push_back(p);
        while (!stack.empty()) {
            const int v = stack.back();
            stack.pop_back();
            order.push_back(v);
            for (auto it = members[v].rbegin(); it != members[v].rend(); ++it) stack.push_back(*it);
        }
    }
    if ((int)order.size() != n) fail(S21_ERR, "ordering lost variables");
}

namespace {

// Maximum transversal: match[c] = row of column c, by shortest augmenting paths (breadth first from each
// unmatched column), starting from the diagonal.  Short paths keep the pivots of a circuit matrix where
// they belong: a voltage source's branch column takes its node's row and that node's column the branch row.
bool transversal(int n, const std::vector<int>& cptr, const std::vector<int>& crow, std::vector<int>& match) {
    match.assign(n, -1);
    std::vector<int> col_of(n, -1);
    for (int c = 0; c < n; c++)
        for (int t = cptr[c]; t < cptr[c + 1]; t++)
            if (crow[t] == c) {
                match[c] = c;
                col_of[c] = c;
                break;
            }
    std::vector<int> seen(n, -1), from_col(n, -1), queue;
    for (int c0 = 0; c0 < n; c0++) {
        if (match[c0] >= 0) continue;
        // BFS over columns; from_col[r] = the column through which row r was reached
        queue.assign(1, c0);
        int free_row = -1;
        for (size_t h = 0; h < queue.size() && free_row < 0; h++) {
            const int c = queue[h];
            for (int t = cptr[c]; t < cptr[c + 1]; t++) {
                const int r = crow[t];
                if (seen[r] == c0) continue;
                seen[r] = c0;
                from_col[r] = c;
                if (col_of[r] < 0) {
                    free_row = r;
                    break;
                }
                queue.push_back(col_of[r]);
            }
        }
        if (free_row < 0) return false;
        // augment: walk back from the free row, each column on the path taking the row it reached
        for (int r = free_row; r >= 0;) {
            const int c = from_col[r];
            const int prev = match[c];
            match[c] = r;
            col_of[r] = c;
            r = prev;  // the row c held, which the previous column on the path reached c through
        }
    }
    return true;
}

}  // namespace

void large_symbolic(int n, const std::vector<int>& rows, const std::vector<int>& cols, const std::vector<double>* vals,
                    LargeSym& S, double relax) {
    S = LargeSym();
    S.n = n;
    if (n <= 0) return;
    const size_t ne = rows.size();
    // ---- columns of A (deduplicated) ----
    auto column_lists = [&](bool nonzero_only, std::vector<int>& cptr, std::vector<int>& crow) {
        std::vector<std::vector<int>> cl(n);
        for (size_t e = 0; e < ne; e++) {
            if (rows[e] < 0 || cols[e] < 0 || rows[e] >= n || cols[e] >= n) fail(S21_INVALID, "matrix entry out of range");
            if (nonzero_only && vals && (*vals)[e] == 0.0) continue;
            cl[cols[e]].push_back(rows[e]);
        }
        cptr.assign(n + 1, 0);
        crow.clear();
        for (int c = 0; c < n; c++) {
            std::sort(cl[c].begin(), cl[c].end());
            cl[c].erase(std::unique(cl[c].begin(), cl[c].end()), cl[c].end());
            crow.insert(crow.end(), cl[c].begin(), cl[c].end());
            cptr[c + 1] = (int)crow.size();
        }
    };
    std::vector<int> cptr, crow, match;
    column_lists(true, cptr, crow);
    if (!transversal(n, cptr, crow, match)) {
        column_lists(false, cptr, crow);
        if (!transversal(n, cptr, crow, match)) fail(S21_SINGULAR, "Singular Matrix");
    } else if (vals) {
        column_lists(false, cptr, crow);  // the full pattern for the ordering
    }
    S.nnz_a = (int64_t)crow.size();
    std::vector<int> col_of_row(n);
    for (int c = 0; c < n; c++) col_of_row[match[c]] = c;
    // ---- symmetric pattern of the row-matched matrix, in column labels ----
    std::vector<std::vector<int>> sym(n);
    for (int c = 0; c < n; c++)
        for (int t = cptr[c]; t < cptr[c + 1]; t++) {
            const int a = col_of_row[crow[t]];
            if (a == c) continue;
            sym[a].push_back(c);
            sym[c].push_back(a);
        }
    std::vector<int> aptr(n + 1, 0), aadj;
    for (int i = 0; i < n; i++) {
        std::sort(sym[i].begin(), sym[i].end());
        sym[i].erase(std::unique(sym[i].begin(), sym[i].end()), sym[i].end());
        aadj.insert(aadj.end(), sym[i].begin(), sym[i].end());
        aptr[i + 1] = (int)aadj.size();
    }
    std::vector<int> order;
    amd_order(n, aptr, aadj, order);
    std::vector<int> pos(n);
    for (int k = 0; k < n; k++) pos[order[k]] = k;
    // ---- elimination tree of the permuted pattern (Liu), then postorder ----
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; k++) {
        const int v = order[k];
        for (int t = aptr[v]; t < aptr[v + 1]; t++) {
            int j = pos[aadj[t]];
            if (j >= k) continue;
            while (j != -1 && j < k) {
                const int nx = anc[j];
                anc[j] = k;
                if (nx == -1) parent[j] = k;
                j = nx;
            }
        }
    }
    std::vector<int> fc(n, -1), ns_(n, -1), post;
    for (int j = n - 1; j >= 0; j--)
        if (parent[j] >= 0) {
            ns_[j] = fc[parent[j]];
            fc[parent[j]] = j;
        }
    post.reserve(n);
    {
        std::vector<int> stk;
        for (int r = 0; r < n; r++) {
            if (parent[r] != -1) continue;
            stk.push_back(r);
            while (!stk.empty()) {
                const int v = stk.back();
                if (fc[v] >= 0) {
                    const int c = fc[v];
                    fc[v] = ns_[c];
                    stk.push_back(c);
                } else {
                    stk.pop_back();
                    post.push_back(v);
                }
            }
        }
    }
    // final order: perm[k] = column label eliminated k-th
    S.perm.resize(n);
    S.prow.resize(n);
    S.icol.resize(n);
    S.irow.resize(n);
    for (int k = 0; k < n; k++) S.perm[k] = order[post[k]];
    for (int k = 0; k < n; k++) {
        S.icol[S.perm[k]] = k;
        S.prow[k] = match[S.perm[k]];
    }
    for (int k = 0; k < n; k++) S.irow[S.prow[k]] = k;
    // internal symmetric adjacency (lower: neighbours below each index)
    std::vector<std::vector<int>> low(n);  // low[i]: j < i adjacent
    for (int v = 0; v < n; v++) {
        const int i = S.icol[v];
        for (int t = aptr[v]; t < aptr[v + 1]; t++) {
            const int j = S.icol[aadj[t]];
            if (j < i) low[i].push_back(j);
        }
    }
    // etree in the new numbering
    std::fill(parent.begin(), parent.end(), -1);
    std::fill(anc.begin(), anc.end(), -1);
    for (int i = 0; i < n; i++)
        for (int j : low[i]) {
            while (j != -1 && j < i) {
                const int nx = anc[j];
                anc[j] = i;
                if (nx == -1) parent[j] = i;
                j = nx;
            }
        }
    // ---- column counts by row subtrees ----
    std::vector<int> cc(n, 1), mk(n, -1);
    for (int i = 0; i < n; i++) {
        mk[i] = i;
        for (int j : low[i])
            for (; mk[j] != i; j = parent[j]) {
                mk[j] = i;
                cc[j]++;
            }
    }
    // ---- fundamental supernodes ----
    std::vector<int> nchild(n, 0);
    for (int j = 0; j < n; j++)
        if (parent[j] >= 0) nchild[parent[j]]++;
    std::vector<int> first{0};
    for (int j = 1; j < n; j++)
        if (!(parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1)) first.push_back(j);
    first.push_back(n);
    int ns = (int)first.size() - 1;
    std::vector<int> sn_of(n);
    for (int s = 0; s < ns; s++)
        for (int j = first[s]; j < first[s + 1]; j++) sn_of[j] = s;
    // row sets of the fundamental supernodes: structure of the first column minus the own columns
    std::vector<std::vector<int>> R(ns);
    std::fill(mk.begin(), mk.end(), -1);
    for (int i = 0; i < n; i++) {
        mk[i] = i;
        for (int j : low[i])
            for (; mk[j] != i; j = parent[j]) {
                mk[j] = i;
                const int s = sn_of[j];
                if (j == first[s] && i >= first[s + 1]) R[s].push_back(i);  // rows come in ascending i
            }
    }
    std::vector<int> sparent(ns, -1);
    for (int s = 0; s < ns; s++) sparent[s] = R[s].empty() ? -1 : sn_of[R[s][0]];
    // ---- relaxed amalgamation: a child merges into its parent when their columns are contiguous and the
    // explicit zeros the merge adds stay within `relax` of the merged supernode's entries ----
    std::vector<int> lead(ns);
    std::iota(lead.begin(), lead.end(), 0);
    std::vector<int64_t> zeros(ns, 0);
    std::vector<int> ncols(ns);
    for (int s = 0; s < ns; s++) ncols[s] = first[s + 1] - first[s];
    std::vector<char> alive(ns, 1);
    // process in postorder: s's parent p has a larger index; s merges into p when s ends where p begins
    for (int s = 0; s < ns; s++) {
        const int p = sparent[s];
        if (p < 0 || relax <= 0.0) continue;
        const int s_last = first[s] + ncols[s];  // s is alive with columns [first[s], s_last)
        if (s_last != first[p]) continue;
        const int64_t kp = ncols[p], ks = ncols[s];
        const int64_t rp = (int64_t)R[p].size();
        // merged front: columns ks + kp, rows below each column: (ks + kp - 1 - j) + rp
        // true nonzeros of s's columns: sum over its columns of the structure below the diagonal
        int64_t true_s = 0;
        for (int64_t j = 0; j < ks; j++) true_s += (ks - 1 - j) + (int64_t)R[s].size() + 0;  // supernode columns
        int64_t merged_s = 0;
        for (int64_t j = 0; j < ks; j++) merged_s += (ks + kp - 1 - j) + rp;
        const int64_t add = merged_s - true_s;
        int64_t total = 0;
        for (int64_t j = 0; j < ks + kp; j++) total += (ks + kp - j) + rp;
        if ((double)(zeros[s] + zeros[p] + add) > relax * (double)total) continue;
        // merge: p takes s's columns (p now starts at first[s])
        zeros[p] += zeros[s] + add;
        first[p] = first[s];
        ncols[p] += ncols[s];
        alive[s] = 0;
        lead[s] = p;
    }
    // compact the surviving supernodes (their first columns ascend with their index)
    std::vector<int> newid(ns, -1);
    int m = 0;
    for (int s = 0; s < ns; s++)
        if (alive[s]) newid[s] = m++;
    S.ns = m;
    S.sn_first.assign(m + 1, n);
    S.sn_parent.assign(m, -1);
    S.r_ptr.assign(m + 1, 0);
    for (int s = 0; s < ns; s++) {
        if (!alive[s]) continue;
        const int t = newid[s];
        S.sn_first[t] = first[s];
        int q = sparent[s];
        while (q >= 0 && !alive[q]) q = lead[q];  // a parent merged into its own parent
        S.sn_parent[t] = q >= 0 ? newid[q] : -1;
        S.r_ptr[t + 1] = S.r_ptr[t] + (int64_t)R[s].size();
        S.r_idx.insert(S.r_idx.end(), R[s].begin(), R[s].end());
    }
    S.sn_first[m] = n;
    for (int t = 0; t < m; t++)
        if (S.sn_parent[t] >= 0 && S.sn_parent[t] <= t) fail(S21_ERR, "supernode tree broken");
    // ---- fronts, children, extend-add maps, levels, statistics ----
    S.front_off.assign(m + 1, 0);
    for (int t = 0; t < m; t++) {
        const int64_t F = (S.sn_first[t + 1] - S.sn_first[t]) + (S.r_ptr[t + 1] - S.r_ptr[t]);
        S.front_off[t + 1] = S.front_off[t] + F * F;
        S.max_front = std::max<int>(S.max_front, (int)F);
        const int64_t k = S.sn_first[t + 1] - S.sn_first[t];
        for (int64_t j = 0; j < k; j++) {
            const double below = (double)(F - j - 1);
            S.flops += below + 2.0 * below * below;
            S.nnz_lu += 2 * (F - j - 1) + 1;
        }
    }
    S.pool = S.front_off[m];
    std::vector<std::vector<int>> ch(m);
    for (int t = 0; t < m; t++)
        if (S.sn_parent[t] >= 0) ch[S.sn_parent[t]].push_back(t);
    S.child_ptr.assign(m + 1, 0);
    for (int t = 0; t < m; t++) {
        S.child.insert(S.child.end(), ch[t].begin(), ch[t].end());
        S.child_ptr[t + 1] = (int)S.child.size();
    }
    S.ext_ptr.assign(m + 1, 0);
    for (int t = 0; t < m; t++) {
        const int p = S.sn_parent[t];
        if (p >= 0) {
            const int f = S.sn_first[p], kp = S.sn_first[p + 1] - S.sn_first[p];
            const int* r0 = S.r_idx.data() + S.r_ptr[p];
            const int* r1 = S.r_idx.data() + S.r_ptr[p + 1];
            for (int64_t q = S.r_ptr[t]; q < S.r_ptr[t + 1]; q++) {
                const int r = S.r_idx[q];
                int at;
                if (r >= f && r < f + kp) {
                    at = r - f;
                } else {
                    const int* it = std::lower_bound(r0, r1, r);
                    if (it == r1 || *it != r) fail(S21_ERR, "child row outside its parent's front");
                    at = kp + (int)(it - r0);
                }
                S.ext_map.push_back(at);
            }
        }
        S.ext_ptr[t + 1] = (int64_t)S.ext_map.size();
    }
    std::vector<int> level(m, 0);
    int maxl = 0;
    for (int t = 0; t < m; t++) {
        for (int c = S.child_ptr[t]; c < S.child_ptr[t + 1]; c++) level[t] = std::max(level[t], level[S.child[c]] + 1);
        maxl = std::max(maxl, level[t]);
    }
    S.lvl_ptr.assign(maxl + 2, 0);
    for (int t = 0; t < m; t++) S.lvl_ptr[level[t] + 1]++;
    for (int l = 0; l <= maxl; l++) S.lvl_ptr[l + 1] += S.lvl_ptr[l];
    S.lvl_sn.assign(m, 0);
    std::vector<int> fill(S.lvl_ptr.begin(), S.lvl_ptr.end() - 1);
    for (int t = 0; t < m; t++) S.lvl_sn[fill[level[t]]++] = t;
}

int64_t large_entry_offset(const LargeSym& S, int a, int b) {
    const int lo = std::min(a, b);
    const int s = (int)(std::upper_bound(S.sn_first.begin(), S.sn_first.end(), lo) - S.sn_first.begin()) - 1;
    const int f = S.sn_first[s], k = S.sn_first[s + 1] - f;
    const int* r0 = S.r_idx.data() + S.r_ptr[s];
    const int* r1 = S.r_idx.data() + S.r_ptr[s + 1];
    const int64_t F = k + (r1 - r0);
    auto local = [&](int x) -> int64_t {
        if (x >= f && x < f + k) return x - f;
        const int* it = std::lower_bound(r0, r1, x);
        if (it == r1 || *it != x) return -1;
        return k + (it - r0);
    };
    const int64_t i = local(a), j = local(b);
    if (i < 0 || j < 0) return -1;
    return S.front_off[s] + j * F + i;
}

void large_factor_host(const LargeSym& S, const std::vector<int>& rows, const std::vector<int>& cols,
                       const std::vector<double>& vals, std::vector<double>& pool) {
    pool.assign((size_t)S.pool, 0.0);
    for (size_t e = 0; e < rows.size(); e++) {
        const int64_t o = large_entry_offset(S, S.irow[rows[e]], S.icol[cols[e]]);
        if (o < 0) fail(S21_ERR, "matrix entry outside the symbolic structure");
        pool[o] += vals[e];
    }
    for (int s = 0; s < S.ns; s++) {
        const int64_t k = S.sn_first[s + 1] - S.sn_first[s];
        const int64_t F = k + (S.r_ptr[s + 1] - S.r_ptr[s]);
        double* A = pool.data() + S.front_off[s];
        for (int c = S.child_ptr[s]; c < S.child_ptr[s + 1]; c++) {
            const int ch = S.child[c];
            const int64_t kc = S.sn_first[ch + 1] - S.sn_first[ch];
            const int64_t Fc = kc + (S.r_ptr[ch + 1] - S.r_ptr[ch]);
            const double* U = pool.data() + S.front_off[ch];
            const int* mp = S.ext_map.data() + S.ext_ptr[ch];
            for (int64_t j = kc; j < Fc; j++)
                for (int64_t i = kc; i < Fc; i++) A[(int64_t)mp[j - kc] * F + mp[i - kc]] += U[j * Fc + i];
        }
        for (int64_t j = 0; j < k; j++) {
            const double piv = A[j * F + j];
            if (piv == 0.0) fail(S21_PIVOT, "Pivot Search Fail");
            for (int64_t i = j + 1; i < F; i++) A[j * F + i] = A[j * F + i] / piv;
            for (int64_t c = j + 1; c < F; c++) {
                const double u = A[c * F + j];
                for (int64_t i = j + 1; i < F; i++) A[c * F + i] = A[c * F + i] - A[j * F + i] * u;
            }
        }
    }
}

void large_solve_host(const LargeSym& S, const std::vector<double>& pool, const std::vector<double>& b,
                      std::vector<double>& x) {
    const int n = S.n;
    std::vector<double> y(n);
    for (int a = 0; a < n; a++) y[a] = b[S.prow[a]];
    std::vector<int64_t> uoff(S.ns + 1, 0);
    for (int s = 0; s < S.ns; s++) uoff[s + 1] = uoff[s] + (S.r_ptr[s + 1] - S.r_ptr[s]);
    std::vector<double> upd((size_t)uoff[S.ns]), w;
    for (int s = 0; s < S.ns; s++) {
        const int f = S.sn_first[s];
        const int64_t k = S.sn_first[s + 1] - f;
        const int64_t F = k + (S.r_ptr[s + 1] - S.r_ptr[s]);
        const double* A = pool.data() + S.front_off[s];
        w.assign((size_t)F, 0.0);
        for (int64_t j = 0; j < k; j++) w[j] = y[f + j];
        for (int c = S.child_ptr[s]; c < S.child_ptr[s + 1]; c++) {
            const int ch = S.child[c];
            const int* mp = S.ext_map.data() + S.ext_ptr[ch];
            for (int64_t q = 0; q < uoff[ch + 1] - uoff[ch]; q++) w[mp[q]] += upd[uoff[ch] + q];
        }
        for (int64_t j = 0; j < k; j++)
            for (int64_t i = j + 1; i < F; i++) w[i] = w[i] - A[j * F + i] * w[j];
        for (int64_t j = 0; j < k; j++) y[f + j] = w[j];
        for (int64_t q = 0; q < F - k; q++) upd[uoff[s] + q] = w[k + q];
    }
    std::vector<double> xi(n, 0.0);
    for (int s = S.ns - 1; s >= 0; s--) {
        const int f = S.sn_first[s];
        const int64_t k = S.sn_first[s + 1] - f;
        const int64_t F = k + (S.r_ptr[s + 1] - S.r_ptr[s]);
        const double* A = pool.data() + S.front_off[s];
        const int* R = S.r_idx.data() + S.r_ptr[s];
        for (int64_t j = k - 1; j >= 0; j--) {
            double acc = y[f + j];
            for (int64_t c = k; c < F; c++) acc = acc - A[c * F + j] * xi[R[c - k]];
            for (int64_t c = j + 1; c < k; c++) acc = acc - A[c * F + j] * xi[f + c];
            xi[f + j] = acc / A[j * F + j];
        }
    }
    x.assign(n, 0.0);
    for (int b2 = 0; b2 < n; b2++) x[S.perm[b2]] = xi[b2];
}

}  // namespace s21
