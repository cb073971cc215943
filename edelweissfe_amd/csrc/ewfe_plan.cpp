// Host-side plan construction for libewfe.so.
//
// Replaces, for the element-assembly path, the reference's per-step set-up:
//   * DofManager's VIJ pattern (edelweissfe/numerics/dofmanager.py:522-557): every element couples all
//     of its dofs with each other, Ke stored column-major in the VIJ slice;
//   * CSRGenerator.__init__ (edelweissfe/numerics/csrgenerator.pyx:47-98): canonical CSR of that COO
//     pattern (rows = dofs, sorted unique columns, explicit zeros kept, int32 indices) and, per VIJ
//     entry, the CSR slot it accumulates into.
// Instead of a slot per VIJ entry (nDof_el^2 per element) the plan keeps, per element and node pair
// (a, b), the offset of node b's 3 columns inside node a's rows: the 3 rows of one node have the same
// columns, and a node's 3 dofs are consecutive, so slot(a,i ; b,k) = indptr[dof(a)+i] + colpos(a,b) + k.
// It also colours the elements (no two elements of one colour share a node) so that the kernels can
// accumulate into CSR without atomics and in a fixed order.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "ewfe_internal.h"

namespace ewfe {

template <typename F>
static void parallel_for(int64_t n, F &&f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto &t : th) t.join();
}

// --- pipeline orders for 8-node elements -----------------------------------------------------------
// The evaluation is cut into chunks of elements. Chunk c's elements are evaluated (element records) after
// those of chunks < c; the row nodes gathered with chunk c are the ones whose incident elements all lie in
// chunks <= c, so a row node's records are complete when its chunk is gathered, whatever the order inside.
//
// boxGen meshes (edelweissfe/generators/boxgen.py:125-185: nodes numbered with the third grid index
// fastest, elements likewise, C3D8 local nodes at the offsets below) get tile-streaming orders: every
// chunk is a slab of element layers along the slowest axis; inside it, the two fast axes are cut into
// tiles of about 16 x 16 columns and each tile is walked layer by layer. Consecutive workgroups then touch
// the same nodes (element pass) and the same element records (gather pass) while they are in L2.
static const int kBoxOff8[8][3] = {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0},
                                   {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}};

static bool detect_boxgen8(const int32_t *conn, int64_t n_el, int64_t n_nodes, const int64_t *node_dof,
                           int64_t E[3])
{
    if (n_el < 1) return false;
    const int64_t n2 = (int64_t)conn[4] - conn[0];   // local node 4 sits one step along axis 1
    const int64_t n12 = (int64_t)conn[3] - conn[0];  // local node 3 sits one step along axis 0
    if (conn[1] - conn[0] != 1 || n2 < 2 || n12 < 2 * n2 || n12 % n2) return false;
    const int64_t n1 = n12 / n2;
    const int64_t e2 = n2 - 1, e1 = n1 - 1;
    if (n_el % (e1 * e2)) return false;
    const int64_t e0 = n_el / (e1 * e2);
    if ((e0 + 1) * n12 != n_nodes) return false;
    for (int64_t n = 0; n < n_nodes; ++n)
        if (node_dof[n] != 3 * n) return false;
    std::atomic<int> bad{0};
    parallel_for(n_el, [&](int64_t b, int64_t e) {
        for (int64_t el = b; el < e && !bad; ++el) {
            const int64_t i0 = el / (e1 * e2), i1 = (el / e2) % e1, i2 = el % e2;
            for (int a = 0; a < 8; ++a) {
                const int64_t nd = (i0 + kBoxOff8[a][0]) * n12 + (i1 + kBoxOff8[a][1]) * n2 + (i2 + kBoxOff8[a][2]);
                if (conn[el * 8 + a] != nd) {
                    bad = 1;
                    break;
                }
            }
        }
    });
    if (bad) return false;
    E[0] = e0;
    E[1] = e1;
    E[2] = e2;
    return true;
}

// [begin, end) ranges of about `t` indices covering [0, n)
static std::vector<std::pair<int64_t, int64_t>> tiles_of(int64_t n, int64_t t)
{
    const int64_t nt = std::max<int64_t>(1, (n + t / 2) / t);
    std::vector<std::pair<int64_t, int64_t>> v;
    for (int64_t k = 0; k < nt; ++k) v.emplace_back(k * n / nt, (k + 1) * n / nt);
    return v;
}

// Element order of an 8-node plan (plan position -> element): structured meshes in tile-streaming order,
// others in element order; also fixes the chunks (plan.n_chunks, plan.pipe_el). The element pass evaluates
// the plan positions in order, so the state and records of consecutive workgroups are contiguous.
static void element_order8(ewfe_plan &plan, const int32_t *conn, const int64_t *node_dof)
{
    const int64_t n_el = plan.n_el;
    int64_t E[3] = {0, 0, 0};
    plan.structured = detect_boxgen8(conn, n_el, plan.n_nodes, node_dof, E);
    // automatic: one chunk (no second stream). Overlapping the element pass with the gather measured slower
    // on MI355X at 128^3 (both passes then compete for the same CU slots). A requested count is capped at
    // one element (structured: one element layer) per chunk.
    int64_t K = plan.n_chunks > 0 ? plan.n_chunks : 1;
    K = std::min<int64_t>(K, plan.structured ? E[0] : n_el);
    plan.n_chunks = (int)K;
    plan.pipe_el.assign(K + 1, 0);
    plan.order.clear();
    plan.order.reserve(n_el);
    if (plan.structured) {
        for (int i = 0; i < 3; ++i) plan.grid_el[i] = E[i];
        const auto et1 = tiles_of(E[1], 16), et2 = tiles_of(E[2], 16);
        for (int64_t c = 0; c < K; ++c) {
            const int64_t l0 = c * E[0] / K, l1 = (c + 1) * E[0] / K;
            for (const auto &t1 : et1)
                for (const auto &t2 : et2)
                    for (int64_t i0 = l0; i0 < l1; ++i0)
                        for (int64_t i1 = t1.first; i1 < t1.second; ++i1)
                            for (int64_t i2 = t2.first; i2 < t2.second; ++i2)
                                plan.order.push_back((i0 * E[1] + i1) * E[2] + i2);
            plan.pipe_el[c + 1] = (int64_t)plan.order.size();
        }
    } else {
        for (int64_t e = 0; e < n_el; ++e) plan.order.push_back(e);
        for (int64_t c = 0; c < K; ++c) plan.pipe_el[c + 1] = (c + 1) * n_el / K;
    }
}

// Row order of an 8-node plan: chunk by chunk, the row nodes whose last incident element is evaluated in the
// chunk; structured meshes walk node tiles layer by layer like the element order.
static int row_order8(ewfe_plan &plan, const std::vector<int32_t> &rownodes, const std::vector<int64_t> &inc_ptr,
                      const std::vector<int64_t> &inc)
{
    const int64_t n_el = plan.n_el, nr = (int64_t)rownodes.size(), K = plan.n_chunks;
    plan.pipe_row.assign(K + 1, 0);
    plan.rorder.clear();
    plan.rorder.reserve(nr);
    std::vector<int32_t> el_chunk(n_el, -1);
    for (int64_t c = 0; c < K; ++c)
        for (int64_t p = plan.pipe_el[c]; p < plan.pipe_el[c + 1]; ++p) el_chunk[plan.order[p]] = (int32_t)c;
    for (int64_t e = 0; e < n_el; ++e)
        if (el_chunk[e] < 0) {
            set_error("internal error: element order does not cover the elements");
            return -1;
        }
    auto last_chunk = [&](int64_t r) {
        int32_t m = 0;
        for (int64_t k = inc_ptr[rownodes[r]]; k < inc_ptr[rownodes[r] + 1]; ++k) m = std::max(m, el_chunk[inc[k]]);
        return m;
    };
    if (plan.structured) {
        const int64_t *E = plan.grid_el;
        const int64_t N1 = E[1] + 1, N2 = E[2] + 1;
        const auto nt1 = tiles_of(N1, 16), nt2 = tiles_of(N2, 16);
        for (int64_t c = 0; c < K; ++c) {
            const int64_t l0 = c * E[0] / K, l1 = (c + 1) * E[0] / K;
            // node layers whose incident element layers (i0 - 1 and i0) end in this chunk
            const int64_t m1 = (c == K - 1) ? l1 + 1 : l1;
            for (const auto &t1 : nt1)
                for (const auto &t2 : nt2)
                    for (int64_t i0 = l0; i0 < m1; ++i0)
                        for (int64_t i1 = t1.first; i1 < t1.second; ++i1)
                            for (int64_t i2 = t2.first; i2 < t2.second; ++i2)
                                plan.rorder.push_back((int32_t)((i0 * N1 + i1) * N2 + i2));
            plan.pipe_row[c + 1] = (int64_t)plan.rorder.size();
        }
        // structured meshes have every node as a row node, row index = node index
        if ((int64_t)plan.rorder.size() != nr) {
            set_error("internal error: structured row order does not cover the row nodes");
            return -1;
        }
    } else {
        std::vector<int32_t> rchunk(nr);
        std::vector<int64_t> cnt(K + 1, 0);
        for (int64_t r = 0; r < nr; ++r) {
            rchunk[r] = last_chunk(r);
            cnt[rchunk[r] + 1]++;
        }
        for (int64_t c = 0; c < K; ++c) cnt[c + 1] += cnt[c];
        for (int64_t c = 0; c <= K; ++c) plan.pipe_row[c] = cnt[c];
        plan.rorder.resize(nr);
        for (int64_t r = 0; r < nr; ++r) plan.rorder[cnt[rchunk[r]]++] = (int32_t)r;
    }
    // every row node once, in the chunk where its last incident element is evaluated
    std::vector<uint8_t> seen(nr, 0);
    for (int64_t c = 0; c < K; ++c)
        for (int64_t q = plan.pipe_row[c]; q < plan.pipe_row[c + 1]; ++q) {
            const int32_t r = plan.rorder[q];
            if (r < 0 || r >= nr || seen[r]) {
                set_error("internal error: row order is not a permutation");
                return -1;
            }
            seen[r] = 1;
            if (last_chunk(r) != c) {
                set_error("internal error: a row node is gathered in another chunk than its last element");
                return -1;
            }
        }
    return 0;
}

int build_plan_host(ewfe_plan &plan, const int32_t *conn, const int64_t *node_dof, std::vector<int32_t> &edof_host,
                    std::vector<int32_t> &colpos_host, GatherHost &gather)
{
    const int npe = plan.npe;
    const int64_t n_el = plan.n_el, n_nodes = plan.n_nodes, n_dof = plan.n_dof;

    // --- validate topology and the dof layout --------------------------------------------------
    for (int64_t e = 0; e < n_el; ++e)
        for (int a = 0; a < npe; ++a) {
            const int64_t nd = conn[e * npe + a];
            if (nd < 0 || nd >= n_nodes) {
                set_error("element " + std::to_string(e) + " references node " + std::to_string(nd) +
                          " outside [0, n_nodes)");
                return -1;
            }
            const int64_t d = node_dof[nd];
            if (d < 0 || d + 3 > n_dof) {
                set_error("node " + std::to_string(nd) + " of element " + std::to_string(e) +
                          " has no valid displacement dofs");
                return -1;
            }
        }
    if (n_dof >= (int64_t)INT32_MAX) {
        set_error("n_dof beyond int32 indexing");
        return -4;
    }
    std::vector<int32_t> node_of_dof(n_dof, -1);
    for (int64_t n = 0; n < n_nodes; ++n) {
        const int64_t d = node_dof[n];
        if (d < 0) continue;
        if (d + 3 > n_dof) {
            set_error("node dof range beyond n_dof");
            return -1;
        }
        for (int c = 0; c < 3; ++c) {
            if (node_of_dof[d + c] != -1) {
                set_error("two nodes share dof " + std::to_string(d + c));
                return -1;
            }
            node_of_dof[d + c] = (int32_t)n;
        }
    }

    // --- node -> element incidence ---------------------------------------------------------------
    std::vector<int64_t> inc_ptr(n_nodes + 1, 0);
    for (int64_t i = 0; i < n_el * npe; ++i) inc_ptr[conn[i] + 1]++;
    for (int64_t n = 0; n < n_nodes; ++n) inc_ptr[n + 1] += inc_ptr[n];
    std::vector<int64_t> inc(n_el * npe);
    {
        std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int64_t e = 0; e < n_el; ++e)
            for (int a = 0; a < npe; ++a) inc[fill[conn[e * npe + a]]++] = e;
    }

    // --- node neighbour lists, sorted by first dof (= column order) ------------------------------
    std::vector<int32_t> deg(n_nodes, 0);
    parallel_for(n_nodes, [&](int64_t b, int64_t e) {
        std::vector<int64_t> tmp;
        for (int64_t n = b; n < e; ++n) {
            tmp.clear();
            for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
                const int64_t el = inc[k];
                for (int a = 0; a < npe; ++a) tmp.push_back(node_dof[conn[el * npe + a]]);
            }
            std::sort(tmp.begin(), tmp.end());
            deg[n] = (int32_t)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
        }
    });
    std::vector<int64_t> nb_ptr(n_nodes + 1, 0);
    for (int64_t n = 0; n < n_nodes; ++n) nb_ptr[n + 1] = nb_ptr[n] + deg[n];
    std::vector<int32_t> nb_dof(nb_ptr[n_nodes]);  // first dof of each neighbour, ascending
    parallel_for(n_nodes, [&](int64_t b, int64_t e) {
        std::vector<int64_t> tmp;
        for (int64_t n = b; n < e; ++n) {
            tmp.clear();
            for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
                const int64_t el = inc[k];
                for (int a = 0; a < npe; ++a) tmp.push_back(node_dof[conn[el * npe + a]]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            for (size_t i = 0; i < tmp.size(); ++i) nb_dof[nb_ptr[n] + i] = (int32_t)tmp[i];
        }
    });

    // --- CSR pattern over dof rows ---------------------------------------------------------------
    plan.indptr.assign(n_dof + 1, 0);
    int64_t acc = 0;
    for (int64_t d = 0; d < n_dof; ++d) {
        plan.indptr[d] = (int32_t)acc;
        const int32_t n = node_of_dof[d];
        acc += (n >= 0) ? 3 * (int64_t)deg[n] : 0;
        if (acc >= (int64_t)INT32_MAX) {
            set_error("nnz beyond int32 CSR indexing");
            return -4;
        }
    }
    plan.indptr[n_dof] = (int32_t)acc;
    plan.nnz = acc;
    plan.indices.resize(acc);
    parallel_for(n_dof, [&](int64_t b, int64_t e) {
        for (int64_t d = b; d < e; ++d) {
            const int32_t n = node_of_dof[d];
            if (n < 0) continue;
            int32_t *out = plan.indices.data() + plan.indptr[d];
            for (int64_t k = nb_ptr[n]; k < nb_ptr[n + 1]; ++k)
                for (int c = 0; c < 3; ++c) *out++ = nb_dof[k] + c;
        }
    });

    // --- greedy element colouring (elements of one colour share no node); 8-node elements need none:
    //     their records are summed per row node by the gather kernel ------------------------------
    std::vector<uint64_t> node_mask(npe == 8 ? 0 : n_nodes, 0);
    std::vector<int8_t> color(n_el, 0);
    int n_colors = 1;
    for (int64_t e = 0; npe != 8 && e < n_el; ++e) {
        uint64_t used = 0;
        for (int a = 0; a < npe; ++a) used |= node_mask[conn[e * npe + a]];
        if (used == ~0ull) {
            set_error("element colouring needs more than 64 colours");
            return -1;
        }
        const int c = __builtin_ctzll(~used);
        color[e] = (int8_t)c;
        n_colors = std::max(n_colors, c + 1);
        for (int a = 0; a < npe; ++a) node_mask[conn[e * npe + a]] |= (1ull << c);
    }
    plan.n_colors = n_colors;
    plan.color_begin.assign(n_colors + 1, 0);
    for (int64_t e = 0; e < n_el; ++e) plan.color_begin[color[e] + 1]++;
    for (int c = 0; c < n_colors; ++c) plan.color_begin[c + 1] += plan.color_begin[c];
    if (npe == 8) {
        element_order8(plan, conn, node_dof);
    } else {
        plan.order.resize(n_el);
        std::vector<int64_t> fill(plan.color_begin.begin(), plan.color_begin.end() - 1);
        for (int64_t e = 0; e < n_el; ++e) plan.order[fill[color[e]]++] = e;
    }


    // --- device images: element dof bases (SoA) and the node-pair slot map (element-major) -------
    edof_host.resize((size_t)npe * n_el);
    const bool need_colpos = npe != 8;  // the 8-node path uses the row-node gather lists instead
    colpos_host.resize(need_colpos ? (size_t)npe * npe * n_el : 0);
    std::atomic<int> bad{0};
    parallel_for(n_el, [&](int64_t b, int64_t e) {
        for (int64_t p = b; p < e; ++p) {
            const int64_t el = plan.order[p];
            for (int a = 0; a < npe; ++a) {
                const int32_t na = conn[el * npe + a];
                edof_host[(size_t)a * n_el + p] = (int32_t)node_dof[na];
                if (!need_colpos) continue;
                const int32_t *lo = nb_dof.data() + nb_ptr[na], *hi = nb_dof.data() + nb_ptr[na + 1];
                for (int bb = 0; bb < npe; ++bb) {
                    const int32_t db = (int32_t)node_dof[conn[el * npe + bb]];
                    const int32_t *it = std::lower_bound(lo, hi, db);
                    if (it == hi || *it != db) bad = 1;
                    colpos_host[((size_t)p * npe + a) * npe + bb] = (int32_t)(3 * (it - lo));
                }
            }
        }
    });
    if (bad) {
        set_error("internal error: slot map lookup failed");
        return -1;
    }

    // --- row nodes and their neighbour blocks (all element kinds) ------------------------------------
    for (int64_t n = 0; n < n_nodes; ++n) plan.max_deg = std::max(plan.max_deg, deg[n]);
    std::vector<int32_t> rownodes;
    for (int64_t d = 0; d < n_dof; d += 1) {
        const int32_t n = node_of_dof[d];
        if (n >= 0 && node_dof[n] == d && inc_ptr[n + 1] > inc_ptr[n]) rownodes.push_back(n);
    }
    const int64_t nr = (int64_t)rownodes.size();
    plan.n_rownodes = nr;
    plan.rows_cover_all_dofs = (3 * nr == n_dof);
    gather.rownode_dof.resize(nr);
    gather.nb_ptr.assign(nr + 1, 0);
    gather.diag_rank.resize(nr);
    for (int64_t r = 0; r < nr; ++r) {
        gather.rownode_dof[r] = (int32_t)node_dof[rownodes[r]];
        gather.nb_ptr[r + 1] = gather.nb_ptr[r] + deg[rownodes[r]];
    }
    plan.n_nbblocks = gather.nb_ptr[nr];
    gather.rn_meta.resize((size_t)4 * nr);
    for (int64_t r = 0; r < nr; ++r) {
        const int32_t d = gather.rownode_dof[r];
        gather.rn_meta[4 * r + 0] = d;
        gather.rn_meta[4 * r + 1] = plan.indptr[d];
        gather.rn_meta[4 * r + 2] = gather.nb_ptr[r];
        gather.rn_meta[4 * r + 3] = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
    }
    gather.nbcol.resize(plan.n_nbblocks);
    parallel_for(nr, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            const int32_t n = rownodes[r];
            for (int32_t k = 0; k < deg[n]; ++k) {
                gather.nbcol[gather.nb_ptr[r] + k] = nb_dof[nb_ptr[n] + k];
                if (nb_dof[nb_ptr[n] + k] == node_dof[n]) gather.diag_rank[r] = k;
            }
        }
    });

    // --- row-node incidence (all element kinds): incident elements in ascending element order ----------
    if (n_el >= ((int64_t)1 << 26)) {
        set_error("too many elements for the incidence packing");
        return -4;
    }
    gather.ninc_ptr.assign(nr + 1, 0);
    for (int64_t r = 0; r < nr; ++r)
        gather.ninc_ptr[r + 1] = gather.ninc_ptr[r] + (int32_t)(inc_ptr[rownodes[r] + 1] - inc_ptr[rownodes[r]]);
    gather.ninc.resize(gather.ninc_ptr[nr]);
    parallel_for(nr, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            const int32_t n = rownodes[r];
            int64_t w = gather.ninc_ptr[r];
            for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k, ++w) {
                const int64_t el = inc[k];
                int a = 0;
                while (a < npe && conn[el * npe + a] != n) ++a;
                gather.ninc[w] = (int32_t)(el * 32 + a);
            }
        }
    });

    // --- row-node gather lists (20-node elements): every row node's incident elements (ascending element order,
    //     code 32 element + local node) and, per neighbour rank m, the (slot, b) pairs of the elements whose node b is
    //     that neighbour (bytes: slot << 5 | b). Taken when every node belongs to at most 8 elements and couples with
    //     at most 128 nodes (hexahedral meshes of the generators); otherwise the plan keeps the colour path only.
    plan.gather20 = false;
    if (npe == 20) {
        bool ok = plan.max_deg <= 128 && n_el < ((int64_t)1 << 26);
        for (int64_t r = 0; ok && r < nr; ++r)
            if (inc_ptr[rownodes[r] + 1] - inc_ptr[rownodes[r]] > 8) ok = false;
        if (ok) {
            std::vector<int64_t> woff(nr + 1, 0);
            for (int64_t r = 0; r < nr; ++r) {
                const int64_t nk = inc_ptr[rownodes[r] + 1] - inc_ptr[rownodes[r]];
                const int64_t dg = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
                woff[r + 1] = woff[r] + nk + (dg + 1 + 20 * nk + 3) / 4;
            }
            if (woff[nr] >= (int64_t)INT32_MAX) {
                set_error("gather lists beyond int32 indexing");
                return -4;
            }
            gather.gdesc.resize((size_t)4 * nr);
            gather.gblob.assign(woff[nr], 0);
            plan.rorder.resize(nr);
            parallel_for(nr, [&](int64_t b0, int64_t e0) {
                std::vector<int> cnt;
                std::vector<uint8_t> rank;
                for (int64_t r = b0; r < e0; ++r) {
                    plan.rorder[r] = (int32_t)r;
                    const int32_t n = rownodes[r];
                    const int32_t nk = (int32_t)(inc_ptr[n + 1] - inc_ptr[n]);
                    const int32_t dg = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
                    const int32_t *lo = nb_dof.data() + nb_ptr[n], *hi = nb_dof.data() + nb_ptr[n + 1];
                    int32_t *g = gather.gdesc.data() + 4 * r;
                    g[0] = gather.rownode_dof[r];
                    g[1] = plan.indptr[gather.rownode_dof[r]];
                    g[2] = (int32_t)woff[r];
                    g[3] = dg | (nk << 8);
                    int32_t *w = gather.gblob.data() + woff[r];
                    uint8_t *cs = reinterpret_cast<uint8_t *>(w + nk), *ent = cs + dg + 1;
                    rank.assign((size_t)20 * nk, 0);
                    cnt.assign(dg + 1, 0);
                    for (int32_t sl = 0; sl < nk; ++sl) {
                        const int64_t el = inc[inc_ptr[n] + sl];  // ascending element order
                        int a = 0;
                        while (a < 20 && conn[el * 20 + a] != n) ++a;
                        w[sl] = (int32_t)(el * 32 + a);
                        for (int bb = 0; bb < 20; ++bb) {
                            const int32_t db = (int32_t)node_dof[conn[el * 20 + bb]];
                            rank[(size_t)20 * sl + bb] = (uint8_t)(std::lower_bound(lo, hi, db) - lo);
                            cnt[rank[(size_t)20 * sl + bb] + 1]++;
                        }
                    }
                    for (int32_t m = 0; m < dg; ++m) cnt[m + 1] += cnt[m];
                    for (int32_t m = 0; m <= dg; ++m) cs[m] = (uint8_t)cnt[m];
                    for (int32_t sl = 0; sl < nk; ++sl)
                        for (int bb = 0; bb < 20; ++bb) {
                            const int m = rank[(size_t)20 * sl + bb];
                            ent[cnt[m]++] = (uint8_t)((sl << 5) | bb);
                        }
                }
            });
            plan.gather20 = true;
        }
    }

    // --- row-node gather lists (8-node elements) ---------------------------------------------------
    if (npe == 8) {
        if (plan.max_deg > 64) {
            set_error("a node couples with more than 64 nodes; the 8-node gather path supports at most 64");
            return -1;
        }
        gather.inc_ptr.assign(nr + 1, 0);
        for (int64_t r = 0; r < nr; ++r)
            gather.inc_ptr[r + 1] = gather.inc_ptr[r] + (int32_t)(inc_ptr[rownodes[r] + 1] - inc_ptr[rownodes[r]]);
        plan.n_inc = gather.inc_ptr[nr];
        gather.inc.resize(plan.n_inc);
        gather.inc_rank.resize((size_t)plan.n_inc * 8);
        parallel_for(nr, [&](int64_t b, int64_t e) {
            for (int64_t r = b; r < e; ++r) {
                const int32_t n = rownodes[r];
                const int32_t *lo = nb_dof.data() + nb_ptr[n], *hi = nb_dof.data() + nb_ptr[n + 1];
                int64_t w = gather.inc_ptr[r];
                for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k, ++w) {
                    const int64_t el = inc[k];  // ascending element order (incidence built in element order)
                    int a = 0;
                    while (a < 8 && conn[el * 8 + a] != n) ++a;
                    gather.inc[w] = (int32_t)(el * 8 + a);
                    for (int bb = 0; bb < 8; ++bb) {
                        const int32_t db = (int32_t)node_dof[conn[el * 8 + bb]];
                        gather.inc_rank[(size_t)w * 8 + bb] = (uint8_t)(std::lower_bound(lo, hi, db) - lo);
                    }
                }
            }
        });
        if (n_el >= ((int64_t)1 << 28)) {
            set_error("too many elements for the gather index packing");
            return -4;
        }
        // contribution lists: for every neighbour rank m of the row node, the (slot, b) pairs whose element node
        // b is that neighbour, in ascending slot (= element) order
        gather.cl_ptr.assign(nr + 1, 0);
        for (int64_t r = 0; r < nr; ++r) {
            const int32_t nk = gather.inc_ptr[r + 1] - gather.inc_ptr[r];
            const int32_t dg = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
            if (8 * nk > 255 || nk > 31) {
                set_error("a node belongs to more than 31 elements; the 8-node gather path supports at most 31");
                return -1;
            }
            gather.cl_ptr[r + 1] = gather.cl_ptr[r] + (dg + 1) + 8 * nk;
        }
        gather.cl.resize(gather.cl_ptr[nr]);
        parallel_for(nr, [&](int64_t b0, int64_t e0) {
            std::vector<int> cnt;
            for (int64_t r = b0; r < e0; ++r) {
                const int32_t nk = gather.inc_ptr[r + 1] - gather.inc_ptr[r];
                const int32_t dg = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
                uint8_t *cs = gather.cl.data() + gather.cl_ptr[r];
                uint8_t *ent = cs + dg + 1;
                cnt.assign(dg + 1, 0);
                for (int32_t sl = 0; sl < nk; ++sl)
                    for (int bb = 0; bb < 8; ++bb) cnt[gather.inc_rank[(size_t)(gather.inc_ptr[r] + sl) * 8 + bb] + 1]++;
                for (int32_t m = 0; m < dg; ++m) cnt[m + 1] += cnt[m];
                for (int32_t m = 0; m <= dg; ++m) cs[m] = (uint8_t)cnt[m];
                for (int32_t sl = 0; sl < nk; ++sl)
                    for (int bb = 0; bb < 8; ++bb) {
                        const int m = gather.inc_rank[(size_t)(gather.inc_ptr[r] + sl) * 8 + bb];
                        ent[cnt[m]++] = (uint8_t)((sl << 3) | bb);
                    }
            }
        });
        const int rc = row_order8(plan, rownodes, inc_ptr, inc);
        if (rc) return rc;
        // descriptors and blobs in gather order
        gather.gdesc.resize((size_t)4 * nr);
        std::vector<int64_t> woff(nr + 1, 0);
        for (int64_t q = 0; q < nr; ++q) {
            const int32_t r = plan.rorder[q];
            const int64_t nk = gather.inc_ptr[r + 1] - gather.inc_ptr[r];
            const int64_t nbytes = gather.cl_ptr[r + 1] - gather.cl_ptr[r];
            woff[q + 1] = woff[q] + nk + (nbytes + 3) / 4;
        }
        if (woff[nr] >= (int64_t)INT32_MAX) {
            set_error("gather lists beyond int32 indexing");
            return -4;
        }
        gather.gblob.assign(woff[nr], 0);
        parallel_for(nr, [&](int64_t b0, int64_t e0) {
            for (int64_t q = b0; q < e0; ++q) {
                const int32_t r = plan.rorder[q];
                const int32_t nk = gather.inc_ptr[r + 1] - gather.inc_ptr[r];
                const int32_t dg = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
                const int32_t d = gather.rownode_dof[r];
                int32_t *g = gather.gdesc.data() + 4 * q;
                g[0] = d;
                g[1] = plan.indptr[d];
                g[2] = (int32_t)woff[q];
                g[3] = dg | (nk << 8);
                int32_t *w = gather.gblob.data() + woff[q];
                for (int32_t k = 0; k < nk; ++k) w[k] = gather.inc[gather.inc_ptr[r] + k];
                std::memcpy(w + nk, gather.cl.data() + gather.cl_ptr[r], gather.cl_ptr[r + 1] - gather.cl_ptr[r]);
            }
        });
    }
    return 0;
}

// --- system plan of several element blocks ------------------------------------------------------------
// The blocks of a model (ew/models/femodel.py keeps every element in one dict; the reference's DofManager and
// CSRGenerator build one VIJ pattern over all of them, dofmanager.py:522-557, csrgenerator.pyx:47-98) each have
// their own plan and pattern over the common dof numbering. The system pattern is the row-wise union of the
// blocks' patterns (sorted unique columns per row: the canonical CSR of the union of their VIJ entries); its
// row-node graph is what the linear solvers read, and maps[b][k] is the position in the system pattern of entry k
// of block b's pattern.
int build_union_host(ewfe_plan &u, const ewfe_plan *const *plans, int n_plans, const int64_t *node_dof,
                     int64_t n_nodes, GatherHost &gather, std::vector<std::vector<int32_t>> &maps)
{
    const int64_t n_dof = u.n_dof;
    u.indptr.assign(n_dof + 1, 0);
    // row lengths of the union
    std::vector<int64_t> len(n_dof, 0);
    parallel_for(n_dof, [&](int64_t b, int64_t e) {
        std::vector<int32_t> tmp;
        for (int64_t r = b; r < e; ++r) {
            tmp.clear();
            for (int k = 0; k < n_plans; ++k) {
                const ewfe_plan &p = *plans[k];
                tmp.insert(tmp.end(), p.indices.begin() + p.indptr[r], p.indices.begin() + p.indptr[r + 1]);
            }
            std::sort(tmp.begin(), tmp.end());
            len[r] = std::unique(tmp.begin(), tmp.end()) - tmp.begin();
        }
    });
    int64_t acc = 0;
    for (int64_t r = 0; r < n_dof; ++r) {
        u.indptr[r] = (int32_t)acc;
        acc += len[r];
        if (acc >= (int64_t)INT32_MAX) {
            set_error("nnz beyond int32 CSR indexing");
            return -4;
        }
    }
    u.indptr[n_dof] = (int32_t)acc;
    u.nnz = acc;
    u.indices.resize(acc);
    maps.assign(n_plans, {});
    for (int k = 0; k < n_plans; ++k) maps[k].resize(plans[k]->nnz);
    parallel_for(n_dof, [&](int64_t b, int64_t e) {
        std::vector<int32_t> tmp;
        for (int64_t r = b; r < e; ++r) {
            tmp.clear();
            for (int k = 0; k < n_plans; ++k) {
                const ewfe_plan &p = *plans[k];
                tmp.insert(tmp.end(), p.indices.begin() + p.indptr[r], p.indices.begin() + p.indptr[r + 1]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            std::copy(tmp.begin(), tmp.end(), u.indices.begin() + u.indptr[r]);
            for (int k = 0; k < n_plans; ++k) {
                const ewfe_plan &p = *plans[k];
                for (int32_t j = p.indptr[r]; j < p.indptr[r + 1]; ++j)
                    maps[k][j] = u.indptr[r] + (int32_t)(std::lower_bound(tmp.begin(), tmp.end(), p.indices[j]) -
                                                         tmp.begin());
            }
        }
    });
    // row nodes: nodes whose first dof has a nonempty row; neighbour blocks = the node's first row in steps of 3
    std::vector<int32_t> rownode_dof;
    for (int64_t n = 0; n < n_nodes; ++n) {
        const int64_t d = node_dof[n];
        if (d < 0) continue;
        if (d + 3 > n_dof) {
            set_error("node dof range beyond n_dof");
            return -1;
        }
        if (u.indptr[d + 1] > u.indptr[d]) rownode_dof.push_back((int32_t)d);
    }
    std::sort(rownode_dof.begin(), rownode_dof.end());
    const int64_t nr = (int64_t)rownode_dof.size();
    u.n_rownodes = nr;
    u.rows_cover_all_dofs = (3 * nr == n_dof);
    gather.rownode_dof = rownode_dof;
    gather.nb_ptr.assign(nr + 1, 0);
    gather.diag_rank.assign(nr, 0);
    for (int64_t r = 0; r < nr; ++r) {
        const int32_t d = rownode_dof[r];
        const int32_t rl = u.indptr[d + 1] - u.indptr[d];
        for (int c = 1; c < 3; ++c)
            if (u.indptr[d + c + 1] - u.indptr[d + c] != rl) {
                set_error("the rows of a node's dofs have different patterns");
                return -1;
            }
        if (rl % 3) {
            set_error("a row's columns are not whole node blocks");
            return -1;
        }
        gather.nb_ptr[r + 1] = gather.nb_ptr[r] + rl / 3;
        u.max_deg = std::max<int32_t>(u.max_deg, rl / 3);
    }
    u.n_nbblocks = gather.nb_ptr[nr];
    gather.nbcol.resize(u.n_nbblocks);
    gather.rn_meta.resize((size_t)4 * nr);
    for (int64_t r = 0; r < nr; ++r) {
        const int32_t d = rownode_dof[r];
        for (int32_t m = 0; m < gather.nb_ptr[r + 1] - gather.nb_ptr[r]; ++m) {
            const int32_t col = u.indices[u.indptr[d] + 3 * m];
            gather.nbcol[gather.nb_ptr[r] + m] = col;
            if (col == d) gather.diag_rank[r] = m;
        }
        gather.rn_meta[4 * r + 0] = d;
        gather.rn_meta[4 * r + 1] = u.indptr[d];
        gather.rn_meta[4 * r + 2] = gather.nb_ptr[r];
        gather.rn_meta[4 * r + 3] = gather.nb_ptr[r + 1] - gather.nb_ptr[r];
    }
    return 0;
}

}  // namespace ewfe
