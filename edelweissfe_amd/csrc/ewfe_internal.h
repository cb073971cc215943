// Internal declarations shared by the plan builder (host) and the kernels of libewfe.so.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

struct ewfe_plan {
    int elem_kind = 0;
    int npe = 0;  // nodes per element
    int nqp = 0;  // quadrature points per element
    int64_t n_el = 0, n_nodes = 0, n_dof = 0, nnz = 0;
    int device = 0;

    // host side
    std::vector<int32_t> indptr;       // [n_dof + 1]
    std::vector<int32_t> indices;      // [nnz]
    std::vector<int64_t> order;        // plan position -> element index
    std::vector<int64_t> color_begin;  // [n_colors + 1] plan-position ranges per colour
    int n_colors = 0;

    // device side (owned)
    int32_t *d_indptr = nullptr;
    int32_t *d_indices = nullptr;
    int32_t *d_edof = nullptr;    // [npe][n_el]   first dof of element node a, plan order
    int32_t *d_colpos = nullptr;  // [n_el][npe*npe] column offset of node b's block inside node a's rows
    int32_t *d_order = nullptr;   // [n_el] plan position -> element index

    // row nodes (nodes with dofs and elements, ascending first dof) and their neighbour blocks: node-block
    // SpMV / block-Jacobi of the solver, and the 8-node gather
    int64_t n_rownodes = 0, n_inc = 0, n_nbblocks = 0;
    int32_t max_deg = 0;
    bool rows_cover_all_dofs = false;  // every dof is one of the 3 dofs of a row node
    int32_t *d_rownode_dof = nullptr;  // [n_rownodes] first dof of each row node, ascending
    int32_t *d_nb_ptr = nullptr;       // [n_rownodes + 1] neighbour-block offsets
    int32_t *d_nbcol = nullptr;        // [n_nbblocks] first dof (column) of each neighbour block, ascending
    int32_t *d_diag_rank = nullptr;    // [n_rownodes] rank of the node itself among its neighbours
    int32_t *d_rn_meta = nullptr;      // [n_rownodes][4] {first dof, CSR offset of its rows, nb_ptr, degree}
    // 8-node gather: per row node (gather order) its incident elements (8 * element + local node index,
    // ascending element) and contribution lists: cstart[deg + 1] then 8 * n_inc entries (slot << 3 | b),
    // grouped by neighbour rank m (entries of m: cstart[m] .. cstart[m + 1]), ascending incident slot inside
    // each group
    int32_t *d_ninc_ptr = nullptr;     // [n_rownodes + 1] incident elements of every row node (all kinds)
    int32_t *d_ninc = nullptr;         // 32 * element + local node, ascending element
    double *d_bweights = nullptr;      // body-force scratch: nodal volume weights per element (lazy)
    double *d_fweights = nullptr;      // surface-load scratch: nodal force vectors per listed element (lazy)
    int64_t fweights_cap = 0;          // listed elements d_fweights holds
    int32_t *d_fslot = nullptr;        // surface-load scratch: element -> list slot, -1 if not listed (lazy)
    int32_t *d_pos_of_el = nullptr;    // element -> plan position (lazy)
    double *d_mtab = nullptr;          // per-element material sets [n_mat][8], or null
    int32_t *d_mid = nullptr;          // [n_el] set of the element at each plan position
    int32_t n_mat = 0;
    int32_t *d_gdesc = nullptr;        // [n_rownodes][4] gather descriptors in gather order
    int32_t *d_gblob = nullptr;        // gather blobs
    // element records (allocated with the plan, FP64-sized, shared by both value precisions): 36 unique 3x3
    // blocks per element, and nodal forces
    void *d_kbuf = nullptr;
    double *d_fe = nullptr;

    // 8-node evaluation in chunks: plan positions [pipe_el[c], pipe_el[c+1]) are evaluated in order (the plan
    // order itself is the evaluation order: tile-streaming for structured meshes), and the row nodes
    // rorder[pipe_row[c], pipe_row[c+1]) -- those whose incident elements are all evaluated by then -- are
    // gathered after them; with more than one chunk on a second stream, beside the next chunk's element pass.
    int n_chunks = 0;  // requested at creation (0 = automatic), the number used after it
    bool structured = false;  // boxGen-ordered structured hexahedral mesh (tile-streaming orders)
    // 20-node elements: records + row-node gather available (every node in at most 8 elements, at most 128
    // neighbours), and the path in use (false: colour-ordered read-modify-write of the CSR values)
    bool gather20 = false;
    bool use_gather20 = false;
    int64_t grid_el[3] = {0, 0, 0};  // elements per axis (slowest first) when structured
    std::vector<int64_t> pipe_el, pipe_row;
    std::vector<int32_t> rorder;  // host copy of the row order (row node index per gather position)
    int32_t *d_rorder = nullptr;
    // system plans of several blocks (ewfe_plan_create_union): no elements of their own; per block the position
    // in this pattern of each entry of the block's pattern
    std::vector<int32_t *> d_umap;
    std::vector<int64_t> umap_len;
    void *aux_stream = nullptr;       // hipStream_t of pipelined plans, created with the plan (default priority)
    std::vector<void *> chunk_events;  // hipEvent_t per chunk + 1 join event
};

namespace ewfe {
void set_error(const std::string &msg);
// EWFE_F32_D64 values: nnz floats, then (from a 256-byte boundary) 9 doubles per row node
inline int64_t d64_offset_bytes(int64_t nnz) { return (4 * nnz + 255) / 256 * 256; }
inline double *d64_blocks(void *vals, int64_t nnz)
{
    return reinterpret_cast<double *>(reinterpret_cast<char *>(vals) + d64_offset_bytes(nnz));
}
inline bool d64_supported(const ewfe_plan *p)
{
    return p->rows_cover_all_dofs && p->d_rn_meta != nullptr && p->d_diag_rank != nullptr && p->d_nbcol != nullptr;
}
// EWFE_F32_D64: the FP64 diagonal blocks rebuilt from the FP32 off-diagonal values (stream-ordered)
int rebuild_diag64(const ewfe_plan *p, void *vals, void *stream);
// EWFE_F32_D64: the FP64 diagonal blocks' rows of Dirichlet dofs made unit rows (stream-ordered)
int dirichlet_diag64(const ewfe_plan *p, const int32_t *rows, int64_t n_rows, void *vals, void *stream);
// Builds pattern, colouring and slot map on the host; fills everything but the device pointers.
// Returns 0 or a negative error code. edof_host ([npe][n_el]) and colpos_host ([n_el][npe*npe]) receive the
// device images.
struct GatherHost {
    std::vector<int32_t> rownode_dof, inc_ptr, inc, nb_ptr, nbcol, diag_rank, rn_meta;
    std::vector<uint8_t> inc_rank, cl;
    std::vector<int32_t> cl_ptr;
    // the gather's inputs in gather order (position q of the plan's rorder): per row node one descriptor
    // {first dof, CSR offset of its rows, word offset of its blob, degree | (incident elements << 8)}, and
    // the blob: incident element codes (8 * element + local node, ascending element) followed by the
    // contribution lists (bytes, padded to a word)
    std::vector<int32_t> gdesc, gblob;
    std::vector<int32_t> ninc_ptr, ninc;  // row node -> incident elements (32 * element + local node)
};
// System plan of several block plans: the row-wise union of their patterns, its row-node graph, and per block
// the position of each of its pattern entries in the union.
int build_union_host(ewfe_plan &u, const ewfe_plan *const *plans, int n_plans, const int64_t *node_dof,
                     int64_t n_nodes, GatherHost &gather, std::vector<std::vector<int32_t>> &maps);
int build_plan_host(ewfe_plan &plan, const int32_t *conn, const int64_t *node_dof, std::vector<int32_t> &edof_host,
                    std::vector<int32_t> &colpos_host, GatherHost &gather);
}  // namespace ewfe
