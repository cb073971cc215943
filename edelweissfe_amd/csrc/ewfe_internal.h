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
    int32_t *d_colpos = nullptr;  // [npe*npe][n_el] column offset of node b's block inside node a's rows
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
    int32_t *d_inc_ptr = nullptr;      // [n_rownodes + 1]
    int32_t *d_inc = nullptr;          // [n_inc] 8 * element + local node index, ascending element
    uint8_t *d_inc_rank = nullptr;     // [n_inc * 8] rank of element node b among the row node's neighbours
    // per row node, at byte offset d_cl_ptr[r]: cstart[deg + 1] then 8 * n_inc entries (slot << 3 | b), grouped by
    // neighbour rank m (entries of m: cstart[m] .. cstart[m + 1]), ascending incident slot inside each group
    int32_t *d_cl_ptr = nullptr;       // [n_rownodes + 1]
    uint8_t *d_cl = nullptr;
    // element records (allocated on first evaluation): 36 unique 3x3 blocks per element, and nodal forces
    void *d_kbuf = nullptr;
    int kbuf_dtype = -1;
    double *d_fe = nullptr;
};

namespace ewfe {
void set_error(const std::string &msg);
// Builds pattern, colouring and slot map on the host; fills everything but the device pointers.
// Returns 0 or a negative error code. edof_host / colpos_host receive the SoA device images.
struct GatherHost {
    std::vector<int32_t> rownode_dof, inc_ptr, inc, nb_ptr, nbcol, diag_rank, rn_meta;
    std::vector<uint8_t> inc_rank, cl;
    std::vector<int32_t> cl_ptr;
};
int build_plan_host(ewfe_plan &plan, const int32_t *conn, const int64_t *node_dof, std::vector<int32_t> &edof_host,
                    std::vector<int32_t> &colpos_host, GatherHost &gather);
}  // namespace ewfe
