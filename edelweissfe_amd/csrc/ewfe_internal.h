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
};

namespace ewfe {
void set_error(const std::string &msg);
// Builds pattern, colouring and slot map on the host; fills everything but the device pointers.
// Returns 0 or a negative error code. edof_host / colpos_host receive the SoA device images.
int build_plan_host(ewfe_plan &plan, const int32_t *conn, const int64_t *node_dof, std::vector<int32_t> &edof_host,
                    std::vector<int32_t> &colpos_host);
}  // namespace ewfe
