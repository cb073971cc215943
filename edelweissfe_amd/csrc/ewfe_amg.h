// Smoothed-aggregation multigrid hierarchy of libewfe.so: the host-built (symbolic) part, shared by the
// pattern builder (ewfe_amg_host.cpp) and the device code (ewfe_amg.hip).
//
// Unknowns are grouped by node: 3 dofs per node, 3x3 blocks. Level 0 is the plan's CSR matrix (row node r owns
// dofs 3r .. 3r+2, its neighbour blocks are the plan's neighbour lists). Level l+1 has one node per aggregate of
// level l. Per level:
//   A_l      node graph (level 0: plan neighbour lists; l >= 1: block-CSR pattern of the Galerkin product)
//   agg      node -> aggregate (= node of level l+1)
//   P        prolongator pattern: row i holds the aggregates of i's neighbours (support of (I - w D^-1 A) P0)
//   pslot    per A_l entry (i, j): the slot of agg(j) inside P row i
//   W        pattern of A_l P: row i = union of the P rows of i's neighbours
//   R        P transposed: per aggregate I, the (row i, P entry t) pairs with P_t in column I, ascending i
// and the pattern of A_{l+1} = P^T W (rows = aggregates).
#pragma once
#include <stdint.h>

#include <vector>

struct ewfe_plan;
struct ewfe_mg;

namespace ewfe {

struct AmgLevelHost {
    int64_t n = 0;                        // nodes
    std::vector<int32_t> aptr, acol;      // node graph / block pattern (sorted columns, includes the diagonal)
    std::vector<int32_t> adiag;           // position of the diagonal block in each row
    // transfer to the next level (empty on the coarsest level)
    std::vector<int32_t> agg;
    int64_t n_agg = 0;
    std::vector<int32_t> pptr, pcol;
    std::vector<uint8_t> pslot;           // [aptr[n]]
    std::vector<int32_t> wptr, wcol;
    std::vector<int32_t> rptr, ri, rt;
    int32_t max_p = 0, max_w = 0, max_a = 0, max_r = 0;
};

struct AmgHost {
    std::vector<AmgLevelHost> levels;  // levels[0] = the fine node graph
};

// Builds the hierarchy from the fine node graph (n nodes, CSR adjacency incl. the diagonal, sorted columns).
// Coarsening stops at <= coarse_max nodes, at max_levels, or when an aggregation does not reduce the size.
// Returns 0 or EWFE_EINVAL (with the error message set) if a row exceeds the device kernels' limits.
int amg_build_host(int64_t n, const std::vector<int32_t> &aptr, const std::vector<int32_t> &acol, int64_t coarse_max,
                   int max_levels, AmgHost &out);

// limits of the device kernels
constexpr int kAmgMaxP = 32;    // prolongator blocks per row
constexpr int kAmgMaxRow = 128;  // blocks per row of A_l (l >= 1) and of W
constexpr int64_t kAmgMaxCoarse = 700;  // nodes of the coarsest level (dense inverse: (3n)^2 doubles)

// V-cycle z = M^-1 r with the hierarchy set up for `vals` (device pointers, stream-ordered on `stream`; kernels
// return at once while *done != 0 when done is not null)
int mg_apply(ewfe_mg *mg, const void *vals, int vals_dtype, const double *r, double *z, const double *done,
             void *stream);
bool mg_ready(const ewfe_mg *mg);
const ewfe_plan *mg_plan(const ewfe_mg *mg);

}  // namespace ewfe
