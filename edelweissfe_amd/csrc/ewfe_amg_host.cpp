// libewfe.so -- host side of the smoothed-aggregation multigrid preconditioner: aggregation and the sparsity
// patterns of P, A P, P^T (A P) on every level. Mesh-only work, done once per plan; the numeric phase
// (ewfe_amg.hip) fills values for each new tangent.
//
// Aggregation follows the classic three-phase scheme of smoothed aggregation (Vanek, Mandel, Brezina 1996; it
// is also what the reference's pyamg-preconditioned GMRES uses, edelweissfe/linsolve/gmres/gmres.py:87-88,
// through pyamg.smoothed_aggregation_solver's standard aggregation): (1) a node whose neighbours are all free
// becomes a root and takes them, (2) each remaining node joins the aggregate of a neighbour, (3) leftovers form
// aggregates with their free neighbours. Every connection counts as strong (the tangent is a 3x3-block
// elasticity-type operator whose couplings are all of one order). On a structured hexahedral mesh phase 1
// tiles the nodes with 3x3x3 blocks, and the Galerkin operator keeps a 27-block stencil.
#include <algorithm>
#include <string>

#include "ewfe.h"
#include "ewfe_amg.h"
#include "ewfe_internal.h"

namespace ewfe {

namespace {

void aggregate(const AmgLevelHost &L, std::vector<int32_t> &agg, int64_t &n_agg)
{
    const int64_t n = L.n;
    agg.assign(n, -1);
    int32_t na = 0;
    for (int64_t i = 0; i < n; ++i) {  // phase 1
        if (agg[i] != -1) continue;
        bool free_nb = true;
        for (int32_t e = L.aptr[i]; e < L.aptr[i + 1] && free_nb; ++e) free_nb = agg[L.acol[e]] == -1;
        if (!free_nb) continue;
        for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e) agg[L.acol[e]] = na;
        ++na;
    }
    std::vector<int32_t> ph1(agg);
    for (int64_t i = 0; i < n; ++i) {  // phase 2
        if (agg[i] != -1) continue;
        for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e)
            if (ph1[L.acol[e]] != -1) {
                agg[i] = ph1[L.acol[e]];
                break;
            }
    }
    for (int64_t i = 0; i < n; ++i) {  // phase 3
        if (agg[i] != -1) continue;
        agg[i] = na;
        for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e)
            if (agg[L.acol[e]] == -1) agg[L.acol[e]] = na;
        ++na;
    }
    n_agg = na;
}

// rows[i] = sorted union over the listed source rows; stamp-based dedup
template <typename F>
void union_rows(int64_t n_rows, int64_t n_cols, std::vector<int32_t> &ptr, std::vector<int32_t> &col, F &&each_row)
{
    std::vector<int64_t> stamp(n_cols, -1);
    std::vector<int32_t> tmp;
    ptr.assign(n_rows + 1, 0);
    col.clear();
    for (int64_t i = 0; i < n_rows; ++i) {
        tmp.clear();
        each_row(i, [&](int32_t c) {
            if (stamp[c] != i) {
                stamp[c] = i;
                tmp.push_back(c);
            }
        });
        std::sort(tmp.begin(), tmp.end());
        col.insert(col.end(), tmp.begin(), tmp.end());
        ptr[i + 1] = (int32_t)col.size();
    }
}

int32_t max_row(const std::vector<int32_t> &ptr)
{
    int32_t m = 0;
    for (size_t i = 0; i + 1 < ptr.size(); ++i) m = std::max(m, ptr[i + 1] - ptr[i]);
    return m;
}

}  // namespace

int amg_build_host(int64_t n, const std::vector<int32_t> &aptr, const std::vector<int32_t> &acol, int64_t coarse_max,
                   int max_levels, AmgHost &out)
{
    out.levels.clear();
    out.levels.emplace_back();
    {
        AmgLevelHost &L0 = out.levels.back();
        L0.n = n;
        L0.aptr = aptr;
        L0.acol = acol;
    }
    for (int lev = 0;; ++lev) {
        AmgLevelHost &L = out.levels[lev];
        L.adiag.assign(L.n, -1);
        for (int64_t i = 0; i < L.n; ++i)
            for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e)
                if (L.acol[e] == i) L.adiag[i] = e;
        for (int64_t i = 0; i < L.n; ++i)
            if (L.adiag[i] < 0) {
                set_error("multigrid: node graph without a diagonal entry");
                return EWFE_EINVAL;
            }
        L.max_a = max_row(L.aptr);
        if (lev > 0 && L.max_a > kAmgMaxRow) {
            set_error("multigrid: coarse operator row with more than " + std::to_string(kAmgMaxRow) + " blocks");
            return EWFE_EINVAL;
        }
        if (L.n <= coarse_max || lev + 1 >= max_levels) break;
        aggregate(L, L.agg, L.n_agg);
        if (L.n_agg >= L.n || L.n_agg < 1) {  // no coarsening
            L.agg.clear();
            L.n_agg = 0;
            break;
        }
        // P pattern and the slot of each A entry's aggregate inside P row i
        union_rows(L.n, L.n_agg, L.pptr, L.pcol, [&](int64_t i, auto &&add) {
            for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e) add(L.agg[L.acol[e]]);
        });
        L.max_p = max_row(L.pptr);
        if (L.max_p > kAmgMaxP) {
            set_error("multigrid: prolongator row with more than " + std::to_string(kAmgMaxP) + " blocks");
            return EWFE_EINVAL;
        }
        L.pslot.assign(L.aptr[L.n], 0);
        for (int64_t i = 0; i < L.n; ++i)
            for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e) {
                const int32_t J = L.agg[L.acol[e]];
                const int32_t *b = L.pcol.data() + L.pptr[i], *x = L.pcol.data() + L.pptr[i + 1];
                L.pslot[e] = (uint8_t)(std::lower_bound(b, x, J) - b);
            }
        // W = A P pattern
        union_rows(L.n, L.n_agg, L.wptr, L.wcol, [&](int64_t i, auto &&add) {
            for (int32_t e = L.aptr[i]; e < L.aptr[i + 1]; ++e) {
                const int32_t j = L.acol[e];
                for (int32_t t = L.pptr[j]; t < L.pptr[j + 1]; ++t) add(L.pcol[t]);
            }
        });
        L.max_w = max_row(L.wptr);
        if (L.max_w > kAmgMaxRow) {
            set_error("multigrid: A P row with more than " + std::to_string(kAmgMaxRow) + " blocks");
            return EWFE_EINVAL;
        }
        // R = P^T as (row, entry) lists per aggregate, ascending row
        L.rptr.assign(L.n_agg + 1, 0);
        for (int32_t t = 0; t < L.pptr[L.n]; ++t) ++L.rptr[L.pcol[t] + 1];
        for (int64_t I = 0; I < L.n_agg; ++I) L.rptr[I + 1] += L.rptr[I];
        L.ri.assign(L.pptr[L.n], 0);
        L.rt.assign(L.pptr[L.n], 0);
        {
            std::vector<int32_t> fill(L.rptr.begin(), L.rptr.end() - 1);
            for (int64_t i = 0; i < L.n; ++i)
                for (int32_t t = L.pptr[i]; t < L.pptr[i + 1]; ++t) {
                    const int32_t I = L.pcol[t];
                    L.ri[fill[I]] = (int32_t)i;
                    L.rt[fill[I]] = t;
                    ++fill[I];
                }
        }
        L.max_r = max_row(L.rptr);
        // next level: pattern of P^T (A P)
        AmgLevelHost next;
        next.n = L.n_agg;
        union_rows(L.n_agg, L.n_agg, next.aptr, next.acol, [&](int64_t I, auto &&add) {
            for (int32_t q = L.rptr[I]; q < L.rptr[I + 1]; ++q) {
                const int32_t i = L.ri[q];
                for (int32_t u = L.wptr[i]; u < L.wptr[i + 1]; ++u) add(L.wcol[u]);
            }
        });
        out.levels.push_back(std::move(next));
    }
    const AmgLevelHost &C = out.levels.back();
    if (C.n > kAmgMaxCoarse) {
        set_error("multigrid: coarsest level has " + std::to_string(C.n) + " nodes (limit " +
                  std::to_string(kAmgMaxCoarse) + ")");
        return EWFE_EINVAL;
    }
    return 0;
}

}  // namespace ewfe
