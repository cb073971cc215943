// libewfe.so -- SpMV and a device-resident preconditioned conjugate-gradient solver for the Newton correction
// K ddU = R (the reference's NIST.linearSolve, edelweissfe/solvers/nonlinearimplicitstatic.py:727-751, which
// hands the row-eliminated matrix of applyDirichletK (:559-593) to a direct solver).
//
// The Dirichlet rows of the matrix are unit rows (ewfe_dirichlet_rows). Starting from x0 = b on those rows and 0
// elsewhere, the initial residual vanishes on them, every search direction keeps them at zero, and the iteration
// is plain CG on the symmetric positive definite free-free block: its solution is the solution of the
// row-eliminated system.
//
// MI355X layout: the 3 CSR rows of a node have one column pattern, made of 3-column blocks (one per neighbour
// node). The node-block SpMV gives each row node a 32-lane group, one lane per neighbour block: one column index
// per 3x3 block (the plan's neighbour list) instead of one per value, and the 3 rows' values are read as
// contiguous runs. Preconditioner: node-block Jacobi (inverse 3x3 diagonal blocks). All scalars stay on the
// device; the host only polls a "done" flag every few iterations. Reductions are blockwise partial sums followed
// by a single-block sum in fixed order, so the iteration history is bitwise reproducible. Meshes whose dofs are
// not all node dofs fall back to a row-wise CSR SpMV with point Jacobi.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "ewfe.h"
#include "ewfe_amg.h"
#include "ewfe_internal.h"

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            ewfe::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                  \
            return EWFE_EHIP;                                                                    \
        }                                                                                        \
    } while (0)

namespace {

constexpr int kBlocks = 4096;  // fixed grid of the row kernels -> fixed number of partial sums
constexpr int kThreads = 256;
constexpr int kLanesPerRow = 16;
constexpr int kRowsPerBlock = kThreads / kLanesPerRow;
constexpr int kLanesPerNode = 32;
constexpr int kNodesPerBlock = kThreads / kLanesPerNode;

enum Scalar { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_BB, S_DONE, S_ITERS, S_FAIL, S_COUNT = 16 };

struct Mat {
    const int32_t *indptr, *indices;                      // CSR
    const int32_t *rn_dof, *nb_ptr, *nbcol, *diag_rank;  // row nodes
    const int4 *rn_meta;                                  // {first dof, CSR offset, nb_ptr, degree} per row node
    const double *d64;  // EWFE_F32_D64: FP64 corrections of the diagonal blocks (9 per row node); else null
    int64_t n, n_rownodes;
    bool blocked;
};

__device__ __forceinline__ double group_sum16(double v)
{
    v += __shfl_xor(v, 8, 16);
    v += __shfl_xor(v, 4, 16);
    v += __shfl_xor(v, 2, 16);
    v += __shfl_xor(v, 1, 16);
    return v;
}

__device__ __forceinline__ double group_sum32(double v)
{
    v += __shfl_xor(v, 16, 32);
    return group_sum16(v);
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials, int stride)
{
    __shared__ double red[NV][kThreads / 64];
    for (int k = 0; k < NV; ++k) {
        double s = v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x / 64] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0;
        for (int w = 0; w < kThreads / 64; ++w) s += red[threadIdx.x][w];
        partials[threadIdx.x * stride + blockIdx.x] = s;
    }
}

// --- row-wise CSR (fallback) -----------------------------------------------------------------------
template <typename VT>
__device__ __forceinline__ double row_dot(const Mat &M, const VT *vals, const double *x, int64_t r, int lane16)
{
    double acc = 0.0;
    const int32_t b = M.indptr[r], e = M.indptr[r + 1];
    for (int32_t k = b + lane16; k < e; k += kLanesPerRow) acc += (double)vals[k] * x[M.indices[k]];
    return group_sum16(acc);
}

// --- node-block rows: (K x)[d..d+2] of row node rn, summed over the 32-lane group ----------------------------
template <typename VT, bool D64>
__device__ __forceinline__ void node_dot(const Mat &M, const VT *__restrict__ vals, const double *__restrict__ x,
                                         int64_t rn, int lane32, double &y0, double &y1, double &y2)
{
    const int4 meta = M.rn_meta[rn];  // one load instead of a chain rn -> dof -> indptr
    const int32_t base = meta.y, nb0 = meta.z, deg = meta.w;
    const int32_t L = 3 * deg;
    double a0 = 0, a1 = 0, a2 = 0;
    // EWFE_F32_D64 (a separate instantiation): E x_node joins the sums, one load per lane issued before the blocks'
    // (lanes 0..8: E's entries, lanes 9..11: the node's x)
    double ev = 0.0;
    if constexpr (D64) {
        if (lane32 < 12) ev = lane32 < 9 ? M.d64[9 * rn + lane32] : x[meta.x + lane32 - 9];
    }
    auto block = [&](int m) {
        const int32_t c = M.nbcol[nb0 + m];
        const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
        const VT *v0 = vals + base + 3 * m;
        const VT *v1 = v0 + L;
        const VT *v2 = v1 + L;
        a0 += (double)v0[0] * x0 + (double)v0[1] * x1 + (double)v0[2] * x2;
        a1 += (double)v1[0] * x0 + (double)v1[1] * x1 + (double)v1[2] * x2;
        a2 += (double)v2[0] * x0 + (double)v2[1] * x1 + (double)v2[2] * x2;
    };
    for (int m = lane32; m < deg; m += kLanesPerNode) block(m);
    if constexpr (D64) {  // lane t < 9: E[t] x[t % 3] into row t / 3
        const double xk = __shfl(ev, 9 + lane32 % 3, 32);
        const double pr = lane32 < 9 ? ev * xk : 0.0;
        if (lane32 < 3)
            a0 += pr;
        else if (lane32 < 6)
            a1 += pr;
        else
            a2 += pr;
    }
    y0 = group_sum32(a0);
    y1 = group_sum32(a1);
    y2 = group_sum32(a2);
}

template <typename VT, bool D64>
__global__ __launch_bounds__(kThreads) void k_spmv(Mat M, const VT *vals, const double *x, double *y)
{
    if (M.blocked) {
        const int lane32 = threadIdx.x % kLanesPerNode;
        for (int64_t rn = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kLanesPerNode; rn < M.n_rownodes;
             rn += (int64_t)gridDim.x * kNodesPerBlock) {
            double y0, y1, y2;
            node_dot<VT, D64>(M, vals, x, rn, lane32, y0, y1, y2);
            if (lane32 == 0) {
                const int32_t d = M.rn_dof[rn];
                y[d] = y0;
                y[d + 1] = y1;
                y[d + 2] = y2;
            }
        }
    } else {
        const int lane16 = threadIdx.x % kLanesPerRow;
        for (int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; r < M.n;
             r += (int64_t)gridDim.x * kRowsPerBlock) {
            const double s = row_dot(M, vals, x, r, lane16);
            if (lane16 == 0) y[r] = s;
        }
    }
}

// preconditioner: node blocks -> inverse 3x3 diagonal block (9 values per node, row-major, pre[9 rn .. 9 rn + 8]);
// fallback -> inverse diagonal (pre[r])
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_pcg_setup(Mat M, const VT *vals, double *pre, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M.n; i += stride) x[i] = 0.0;
    if (M.blocked) {
        for (int64_t rn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rn < M.n_rownodes; rn += stride) {
            const int32_t d = M.rn_dof[rn];
            const int32_t base = M.indptr[d];
            const int32_t L = M.indptr[d + 1] - base;
            const int32_t off = base + 3 * M.diag_rank[rn];
            double B[3][3];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k)
                    B[i][k] = (double)vals[off + i * L + k] + (M.d64 ? M.d64[9 * rn + 3 * i + k] : 0.0);
            const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                               B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                               B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
            double *Q = pre + 9 * rn;
            if (det != 0.0 && det == det) {
                const double r = 1.0 / det;
                Q[0] = (B[1][1] * B[2][2] - B[1][2] * B[2][1]) * r;
                Q[1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * r;
                Q[2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * r;
                Q[3] = (B[1][2] * B[2][0] - B[1][0] * B[2][2]) * r;
                Q[4] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * r;
                Q[5] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * r;
                Q[6] = (B[1][0] * B[2][1] - B[1][1] * B[2][0]) * r;
                Q[7] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * r;
                Q[8] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * r;
            } else {  // singular block: point Jacobi on it
                for (int t = 0; t < 9; ++t) Q[t] = 0.0;
                for (int i = 0; i < 3; ++i) Q[4 * i] = (B[i][i] != 0.0) ? 1.0 / B[i][i] : 1.0;
            }
        }
    } else {
        for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < M.n; r += stride) {
            double dg = 1.0;
            for (int32_t k = M.indptr[r]; k < M.indptr[r + 1]; ++k)
                if (M.indices[k] == r) {
                    dg = (double)vals[k];
                    break;
                }
            pre[r] = (dg != 0.0) ? 1.0 / dg : 1.0;
        }
    }
}

__global__ void k_fill0(int64_t n, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = 0.0;
}

__global__ void k_set_rows(const int32_t *rows, int64_t n_rows, const double *b, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride) x[rows[i]] = b[rows[i]];
}

__device__ __forceinline__ void apply_block(const double *Q, const double r0, const double r1, const double r2,
                                            double &z0, double &z1, double &z2)
{
    z0 = Q[0] * r0 + Q[1] * r1 + Q[2] * r2;
    z1 = Q[3] * r0 + Q[4] * r1 + Q[5] * r2;
    z2 = Q[6] * r0 + Q[7] * r1 + Q[8] * r2;
}

// r = b - K x, z = M^-1 r, p = z, partials of (r.z, r.r, b.b)
template <typename VT, bool D64>
__global__ __launch_bounds__(kThreads) void k_pcg_residual(Mat M, const VT *vals, const double *b, const double *x,
                                                           const double *pre, double *r, double *z, double *p,
                                                           double *partials)
{
    double acc[3] = {0, 0, 0};
    if (M.blocked) {
        const int lane32 = threadIdx.x % kLanesPerNode;
        for (int64_t rn = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kLanesPerNode; rn < M.n_rownodes;
             rn += (int64_t)gridDim.x * kNodesPerBlock) {
            double k0, k1, k2;
            node_dot<VT, D64>(M, vals, x, rn, lane32, k0, k1, k2);
            if (lane32 == 0) {
                const int32_t d = M.rn_dof[rn];
                const double b0 = b[d], b1 = b[d + 1], b2 = b[d + 2];
                const double r0 = b0 - k0, r1 = b1 - k1, r2 = b2 - k2;
                double z0, z1, z2;
                apply_block(pre + 9 * rn, r0, r1, r2, z0, z1, z2);
                r[d] = r0, r[d + 1] = r1, r[d + 2] = r2;
                z[d] = z0, z[d + 1] = z1, z[d + 2] = z2;
                p[d] = z0, p[d + 1] = z1, p[d + 2] = z2;
                acc[0] += r0 * z0 + r1 * z1 + r2 * z2;
                acc[1] += r0 * r0 + r1 * r1 + r2 * r2;
                acc[2] += b0 * b0 + b1 * b1 + b2 * b2;
            }
        }
    } else {
        const int lane16 = threadIdx.x % kLanesPerRow;
        for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < M.n;
             row += (int64_t)gridDim.x * kRowsPerBlock) {
            const double kx = row_dot(M, vals, x, row, lane16);
            if (lane16 == 0) {
                const double bi = b[row];
                const double ri = bi - kx;
                const double zi = pre[row] * ri;
                r[row] = ri;
                z[row] = zi;
                p[row] = zi;
                acc[0] += ri * zi;
                acc[1] += ri * ri;
                acc[2] += bi * bi;
            }
        }
    }
    block_partials<3>(acc, partials, kBlocks);
}

// Ap = K p, partial of p.Ap
template <typename VT, bool D64>
__global__ __launch_bounds__(kThreads) void k_pcg_spmv(Mat M, const VT *vals, const double *p, double *Ap,
                                                       double *partials, const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    double acc[1] = {0};
    if (M.blocked) {
        const int lane32 = threadIdx.x % kLanesPerNode;
        for (int64_t rn = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kLanesPerNode; rn < M.n_rownodes;
             rn += (int64_t)gridDim.x * kNodesPerBlock) {
            double y0, y1, y2;
            node_dot<VT, D64>(M, vals, p, rn, lane32, y0, y1, y2);
            if (lane32 < 3) {  // lane c: component c (every lane of the group holds the sums)
                const int32_t d = M.rn_dof[rn] + lane32;
                const double y = lane32 == 0 ? y0 : (lane32 == 1 ? y1 : y2);
                Ap[d] = y;
                acc[0] += y * p[d];
            }
        }
    } else {
        const int lane16 = threadIdx.x % kLanesPerRow;
        for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < M.n;
             row += (int64_t)gridDim.x * kRowsPerBlock) {
            const double s = row_dot(M, vals, p, row, lane16);
            if (lane16 == 0) {
                Ap[row] = s;
                acc[0] += s * p[row];
            }
        }
    }
    block_partials<1>(acc, partials, kBlocks);
}

// x += alpha p, r -= alpha Ap, z = M^-1 r; partials of (r.z, r.r)
__global__ __launch_bounds__(kThreads) void k_pcg_update(Mat M, double *x, double *r, double *z, const double *p,
                                                         const double *Ap, const double *pre, double *partials,
                                                         const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double alpha = scal[S_ALPHA];
    double acc[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (M.blocked) {
        for (int64_t rn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rn < M.n_rownodes; rn += stride) {
            const int32_t d = M.rn_dof[rn];
            double rr[3];
            for (int c = 0; c < 3; ++c) {
                x[d + c] += alpha * p[d + c];
                rr[c] = r[d + c] - alpha * Ap[d + c];
                r[d + c] = rr[c];
            }
            double z0, z1, z2;
            apply_block(pre + 9 * rn, rr[0], rr[1], rr[2], z0, z1, z2);
            z[d] = z0, z[d + 1] = z1, z[d + 2] = z2;
            acc[0] += rr[0] * z0 + rr[1] * z1 + rr[2] * z2;
            acc[1] += rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M.n; i += stride) {
            x[i] += alpha * p[i];
            const double ri = r[i] - alpha * Ap[i];
            const double zi = pre[i] * ri;
            r[i] = ri;
            z[i] = zi;
            acc[0] += ri * zi;
            acc[1] += ri * ri;
        }
    }
    block_partials<2>(acc, partials, kBlocks);
}

__global__ __launch_bounds__(kThreads) void k_pcg_pupdate(int64_t n, double *p, const double *z, const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double beta = scal[S_BETA];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = z[i] + beta * p[i];
}

__device__ double sum_partials(const double *partials)
{
    __shared__ double red[1024];
    double s = 0;
    for (int i = threadIdx.x; i < kBlocks; i += blockDim.x) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// mode 0: after residual (rz, rr, bb); mode 1: alpha from p.Ap; mode 2: beta from (rz_new, rr)
__global__ __launch_bounds__(1024) void k_pcg_reduce(int mode, const double *partials, double *scal, double rtol2,
                                                     int maxit)
{
    if (mode != 0 && scal[S_DONE] != 0.0) return;
    if (mode == 0) {
        const double rz = sum_partials(partials);
        const double rr = sum_partials(partials + kBlocks);
        const double bb = sum_partials(partials + 2 * kBlocks);
        if (threadIdx.x == 0) {
            scal[S_RZ] = rz;
            scal[S_RR] = rr;
            scal[S_BB] = bb;
            scal[S_ITERS] = 0;
            scal[S_FAIL] = 0;
            scal[S_DONE] = (rr <= rtol2 * bb || rr == 0.0) ? 1.0 : 0.0;
        }
    } else if (mode == 1) {
        const double pAp = sum_partials(partials);
        if (threadIdx.x == 0) {
            scal[S_PAP] = pAp;
            if (!(pAp > 0.0)) {  // not positive definite on the free block (or NaN)
                scal[S_FAIL] = 1;
                scal[S_DONE] = 1;
            } else {
                scal[S_ALPHA] = scal[S_RZ] / pAp;
            }
        }
    } else {
        const double rz = sum_partials(partials);
        const double rr = sum_partials(partials + kBlocks);
        if (threadIdx.x == 0) {
            scal[S_BETA] = rz / scal[S_RZ];
            scal[S_RZ] = rz;
            scal[S_RR] = rr;
            scal[S_ITERS] += 1;
            if (rr <= rtol2 * scal[S_BB] || scal[S_ITERS] >= maxit || !(rr == rr)) scal[S_DONE] = 1;
        }
    }
}

// multigrid-preconditioned variant: r = b - K x with partials of (r.r, b.b) at partials + kBlocks
template <typename VT, bool D64>
__global__ __launch_bounds__(kThreads) void k_pcg_residual_nz(Mat M, const VT *vals, const double *b,
                                                              const double *x, double *r, double *partials)
{
    double acc[2] = {0, 0};
    const int lane32 = threadIdx.x % kLanesPerNode;
    for (int64_t rn = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kLanesPerNode; rn < M.n_rownodes;
         rn += (int64_t)gridDim.x * kNodesPerBlock) {
        double k0, k1, k2;
        node_dot<VT, D64>(M, vals, x, rn, lane32, k0, k1, k2);
        if (lane32 < 3) {
            const int32_t d = M.rn_dof[rn] + lane32;
            const double bc = b[d], rc = bc - (lane32 == 0 ? k0 : (lane32 == 1 ? k1 : k2));
            r[d] = rc;
            acc[0] += rc * rc;
            acc[1] += bc * bc;
        }
    }
    block_partials<2>(acc, partials + kBlocks, kBlocks);
}

// x += alpha p, r -= alpha Ap; partial of r.r at partials + kBlocks
__global__ __launch_bounds__(kThreads) void k_pcg_update_nz(int64_t n, double *x, double *r, const double *p,
                                                            const double *Ap, double *partials, const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double alpha = scal[S_ALPHA];
    double acc[1] = {0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, partials + kBlocks, kBlocks);
}

// partial of r.z at partials (scal may be null: no early exit)
__global__ __launch_bounds__(kThreads) void k_pcg_dot(int64_t n, const double *r, const double *z, double *partials,
                                                      const double *scal)
{
    if (scal && scal[S_DONE] != 0.0) return;
    double acc[1] = {0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) acc[0] += r[i] * z[i];
    block_partials<1>(acc, partials, kBlocks);
}

Mat make_mat(const ewfe_plan *p, const double *d64 = nullptr)
{
    Mat M;
    M.d64 = d64;
    M.indptr = p->d_indptr;
    M.indices = p->d_indices;
    M.rn_dof = p->d_rownode_dof;
    M.nb_ptr = p->d_nb_ptr;
    M.nbcol = p->d_nbcol;
    M.diag_rank = p->d_diag_rank;
    M.rn_meta = reinterpret_cast<const int4 *>(p->d_rn_meta);
    M.n = p->n_dof;
    M.n_rownodes = p->n_rownodes;
    M.blocked = p->rows_cover_all_dofs && p->d_nbcol != nullptr && p->d_rn_meta != nullptr;
    return M;
}

template <typename VT, bool D64>
int pcg(const ewfe_plan *p, const VT *vals, const double *d64, const double *b, double *x, const int32_t *rows,
        int64_t n_rows, double rtol, int maxit, int check_every, double *ws, int *iters_out, double *relres_out,
        hipStream_t s)
{
    const Mat M = make_mat(p, d64);
    const int64_t n = p->n_dof;
    double *r = ws, *z = ws + n, *pp = ws + 2 * n, *Ap = ws + 3 * n, *pre = ws + 4 * n;
    double *partials = pre + 3 * n, *scal = partials + 3 * kBlocks;
    const double rtol2 = rtol * rtol;
    hipLaunchKernelGGL(k_pcg_setup<VT>, dim3(kBlocks), dim3(kThreads), 0, s, M, vals, pre, x);
    if (n_rows > 0) hipLaunchKernelGGL(k_set_rows, dim3(256), dim3(256), 0, s, rows, n_rows, b, x);
    hipLaunchKernelGGL((k_pcg_residual<VT, D64>), dim3(kBlocks), dim3(kThreads), 0, s, M, vals, b, (const double *)x,
                       (const double *)pre, r, z, pp, partials);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 0, (const double *)partials, scal, rtol2, maxit);
    HIP_CHECK(hipGetLastError());
    double host_scal[S_COUNT];
    int it = 0;
    for (;;) {
        HIP_CHECK(hipMemcpyAsync(host_scal, scal, sizeof(host_scal), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (host_scal[S_DONE] != 0.0 || it >= maxit) break;
        for (int k = 0; k < check_every && it < maxit; ++k, ++it) {
            hipLaunchKernelGGL((k_pcg_spmv<VT, D64>), dim3(kBlocks), dim3(kThreads), 0, s, M, vals, (const double *)pp, Ap,
                               partials, (const double *)scal);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 1, (const double *)partials, scal, rtol2,
                               maxit);
            hipLaunchKernelGGL(k_pcg_update, dim3(kBlocks), dim3(kThreads), 0, s, M, x, r, z, (const double *)pp,
                               (const double *)Ap, (const double *)pre, partials, (const double *)scal);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 2, (const double *)partials, scal, rtol2,
                               maxit);
            hipLaunchKernelGGL(k_pcg_pupdate, dim3(kBlocks), dim3(kThreads), 0, s, n, pp, (const double *)z,
                               (const double *)scal);
        }
        HIP_CHECK(hipGetLastError());
    }
    if (iters_out) *iters_out = (int)host_scal[S_ITERS];
    if (relres_out)
        *relres_out = host_scal[S_BB] > 0 ? std::sqrt(host_scal[S_RR] / host_scal[S_BB]) : std::sqrt(host_scal[S_RR]);
    if (host_scal[S_FAIL] != 0.0) {
        ewfe::set_error("ewfe_pcg: matrix is not positive definite on the free dofs (p.Ap <= 0)");
        return EWFE_EINVAL;
    }
    return 0;
}

template <typename VT, bool D64>
int pcg_mg(ewfe_mg *mg, const ewfe_plan *p, int vals_dtype, const VT *vals, const double *d64, const double *b,
           double *x, const int32_t *rows, int64_t n_rows, double rtol, int maxit, int check_every, double *ws,
           int *iters_out, double *relres_out, hipStream_t s)
{
    const Mat M = make_mat(p, d64);
    const int64_t n = p->n_dof;
    double *r = ws, *z = ws + n, *pp = ws + 2 * n, *Ap = ws + 3 * n;
    double *partials = ws + 7 * n, *scal = partials + 3 * kBlocks;
    const double rtol2 = rtol * rtol;
    void *sv = reinterpret_cast<void *>(s);
    hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, s, n, x);
    if (n_rows > 0) hipLaunchKernelGGL(k_set_rows, dim3(256), dim3(256), 0, s, rows, n_rows, b, x);
    hipLaunchKernelGGL((k_pcg_residual_nz<VT, D64>), dim3(kBlocks), dim3(kThreads), 0, s, M, vals, b, (const double *)x, r,
                       partials);
    int rc = ewfe::mg_apply(mg, vals, vals_dtype, r, z, nullptr, sv);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pcg_dot, dim3(kBlocks), dim3(kThreads), 0, s, n, (const double *)r, (const double *)z,
                       partials, (const double *)nullptr);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 0, (const double *)partials, scal, rtol2, maxit);
    HIP_CHECK(hipMemcpyAsync(pp, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipGetLastError());
    double host_scal[S_COUNT];
    int it = 0;
    for (;;) {
        HIP_CHECK(hipMemcpyAsync(host_scal, scal, sizeof(host_scal), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (host_scal[S_DONE] != 0.0 || it >= maxit) break;
        for (int k = 0; k < check_every && it < maxit; ++k, ++it) {
            hipLaunchKernelGGL((k_pcg_spmv<VT, D64>), dim3(kBlocks), dim3(kThreads), 0, s, M, vals, (const double *)pp, Ap,
                               partials, (const double *)scal);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 1, (const double *)partials, scal, rtol2,
                               maxit);
            hipLaunchKernelGGL(k_pcg_update_nz, dim3(kBlocks), dim3(kThreads), 0, s, n, x, r, (const double *)pp,
                               (const double *)Ap, partials, (const double *)scal);
            rc = ewfe::mg_apply(mg, vals, vals_dtype, r, z, scal + S_DONE, sv);
            if (rc) return rc;
            hipLaunchKernelGGL(k_pcg_dot, dim3(kBlocks), dim3(kThreads), 0, s, n, (const double *)r,
                               (const double *)z, partials, (const double *)scal);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 2, (const double *)partials, scal, rtol2,
                               maxit);
            hipLaunchKernelGGL(k_pcg_pupdate, dim3(kBlocks), dim3(kThreads), 0, s, n, pp, (const double *)z,
                               (const double *)scal);
        }
        HIP_CHECK(hipGetLastError());
    }
    if (iters_out) *iters_out = (int)host_scal[S_ITERS];
    if (relres_out)
        *relres_out = host_scal[S_BB] > 0 ? std::sqrt(host_scal[S_RR] / host_scal[S_BB]) : std::sqrt(host_scal[S_RR]);
    if (host_scal[S_FAIL] != 0.0) {
        ewfe::set_error("ewfe_pcg_mg: matrix or preconditioner is not positive definite on the free dofs");
        return EWFE_EINVAL;
    }
    return 0;
}

// Values of a vals_dtype: FP64 (f64 true), FP32, or FP32 with FP64 diagonal blocks (d64 set); false for another
// dtype, or for EWFE_F32_D64 on a plan without row nodes covering every dof
bool parse_vals(const ewfe_plan *p, const void *vals, int vals_dtype, bool &f64, const double *&d64)
{
    f64 = vals_dtype == EWFE_F64;
    d64 = nullptr;
    if (vals_dtype == EWFE_F64 || vals_dtype == EWFE_F32) return true;
    if (vals_dtype != EWFE_F32_D64 || !p || !ewfe::d64_supported(p)) return false;
    d64 = ewfe::d64_blocks(const_cast<void *>(vals), p->nnz);
    return true;
}

// EWFE_F32_D64: E_r = -(sum of row node r's off-diagonal blocks) - (its diagonal block), the FP32 values summed in
// FP64 (32 lanes per node, a fixed lane assignment and butterfly: bitwise reproducible; such sums are exact unless the
// values' magnitudes spread over more than ~24 binades, so any order gives the same corrections). K_e of every element maps a
// rigid translation to zero (the shape-function gradients of an element sum to zero), so the FP64 matrix has
// diagonal blocks equal to minus the sums of the off-diagonal ones up to its own rounding; with E added to the
// rounded diagonal blocks every row of the stored matrix sums to zero again.
__global__ __launch_bounds__(kThreads) void k_rebuild_diag64(Mat M, const float *__restrict__ vals,
                                                             double *__restrict__ d64)
{
    const int lane32 = threadIdx.x % kLanesPerNode;
    for (int64_t rn = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kLanesPerNode; rn < M.n_rownodes;
         rn += (int64_t)gridDim.x * kNodesPerBlock) {
        const int4 meta = M.rn_meta[rn];
        const int32_t base = meta.y, deg = meta.w;
        const int32_t L = 3 * deg;
        const int32_t dr = M.diag_rank[rn];
        double a[9], dg[9];
        for (int t = 0; t < 9; ++t) a[t] = dg[t] = 0.0;
        for (int m = lane32; m < deg; m += kLanesPerNode) {
            const float *v = vals + base + 3 * m;
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) {
                    const double x = (double)v[i * L + k];
                    if (m == dr)
                        dg[3 * i + k] = x;
                    else
                        a[3 * i + k] += x;
                }
        }
        double out = 0.0;
        for (int t = 0; t < 9; ++t) {
            const double x = group_sum32(a[t]);
            const double d = group_sum32(dg[t]);  // the one lane holding the block: exact
            if (lane32 == t) out = -x - d;
        }
        if (lane32 < 9) d64[9 * rn + lane32] = out;
    }
}

// EWFE_F32_D64 Dirichlet rows: row c of the diagonal correction of the row's node becomes zero, leaving the unit row
// of the FP32 values (row nodes cover every dof, so node r holds dofs 3 r .. 3 r + 2)
__global__ void k_dirichlet_diag64(const int32_t *__restrict__ rows, int64_t n_rows, double *__restrict__ d64)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    const int32_t row = rows[i];
    double *D = d64 + 9 * (int64_t)(row / 3) + 3 * (row % 3);
    for (int k = 0; k < 3; ++k) D[k] = 0.0;
}

}  // namespace

int ewfe::rebuild_diag64(const ewfe_plan *p, void *vals, void *stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (p->n_rownodes == 0) return 0;
    hipLaunchKernelGGL(k_rebuild_diag64, dim3(kBlocks), dim3(kThreads), 0, s, make_mat(p), (const float *)vals,
                       ewfe::d64_blocks(vals, p->nnz));
    HIP_CHECK(hipGetLastError());
    return 0;
}

int ewfe::dirichlet_diag64(const ewfe_plan *p, const int32_t *rows, int64_t n_rows, void *vals, void *stream)
{
    if (n_rows <= 0) return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_dirichlet_diag64, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, s, rows, n_rows,
                       ewfe::d64_blocks(vals, p->nnz));
    HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int64_t ewfe_vals_bytes(const ewfe_plan *p, int vals_dtype)
{
    if (!p) {
        ewfe::set_error("ewfe_vals_bytes: null plan");
        return EWFE_EINVAL;
    }
    if (vals_dtype == EWFE_F64) return 8 * p->nnz;
    if (vals_dtype == EWFE_F32 || vals_dtype == EWFE_F32_MFMA) return 4 * p->nnz;
    // a size only: answered from the host description, also for plans without device data
    if (vals_dtype == EWFE_F32_D64 && p->rows_cover_all_dofs) return ewfe::d64_offset_bytes(p->nnz) + 72 * p->n_rownodes;
    ewfe::set_error("ewfe_vals_bytes: dtype must be EWFE_F64, EWFE_F32, EWFE_F32_MFMA or EWFE_F32_D64 (the latter on a "
                    "plan whose row nodes cover every dof)");
    return EWFE_EINVAL;
}

extern "C" int ewfe_pcg_mg(ewfe_mg *mg, const void *vals, int vals_dtype, const double *b, double *x,
                           const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                           void *workspace, int *iters_out, double *relres_out, void *stream)
{
    const ewfe_plan *p = ewfe::mg_plan(mg);
    bool f64;
    const double *d64;
    if (!p || !vals || !b || !x || !workspace || maxit < 0 || check_every <= 0 ||
        !parse_vals(p, vals, vals_dtype, f64, d64) || !(rtol >= 0.0) || (n_rows > 0 && !rows_dev)) {
        ewfe::set_error("ewfe_pcg_mg: invalid arguments");
        return EWFE_EINVAL;
    }
    if (!ewfe::mg_ready(mg)) {
        ewfe::set_error("ewfe_pcg_mg: the multigrid hierarchy is not set up (ewfe_mg_setup)");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *ws = reinterpret_cast<double *>(workspace);
    if (f64)
        return pcg_mg<double, false>(mg, p, EWFE_F64, (const double *)vals, nullptr, b, x, rows_dev, n_rows, rtol,
                                     maxit, check_every, ws, iters_out, relres_out, s);
    // the V-cycle reads the FP32 values (the hierarchy is set up from them); the CG operator adds the corrections
    if (d64)
        return pcg_mg<float, true>(mg, p, EWFE_F32, (const float *)vals, d64, b, x, rows_dev, n_rows, rtol, maxit,
                                   check_every, ws, iters_out, relres_out, s);
    return pcg_mg<float, false>(mg, p, EWFE_F32, (const float *)vals, nullptr, b, x, rows_dev, n_rows, rtol, maxit,
                                check_every, ws, iters_out, relres_out, s);
}

extern "C" int64_t ewfe_pcg_workspace_size(const ewfe_plan *p)
{
    if (!p) return 0;
    return (int64_t)sizeof(double) * (7 * p->n_dof + 3 * kBlocks + S_COUNT);
}

extern "C" int ewfe_spmv(const ewfe_plan *p, const void *vals, int vals_dtype, const double *x, double *y,
                         void *stream)
{
    bool f64;
    const double *d64;
    if (!p || !p->d_indptr || !vals || !x || !y || !parse_vals(p, vals, vals_dtype, f64, d64)) {
        ewfe::set_error("ewfe_spmv: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const Mat M = make_mat(p, d64);
    if (f64)
        hipLaunchKernelGGL((k_spmv<double, false>), dim3(kBlocks), dim3(kThreads), 0, s, M, (const double *)vals, x,
                           y);
    else if (d64)
        hipLaunchKernelGGL((k_spmv<float, true>), dim3(kBlocks), dim3(kThreads), 0, s, M, (const float *)vals, x, y);
    else
        hipLaunchKernelGGL((k_spmv<float, false>), dim3(kBlocks), dim3(kThreads), 0, s, M, (const float *)vals, x, y);
    HIP_CHECK(hipGetLastError());
    return 0;
}

// r = b - K x with every row's sum carried in twice the working precision: the compensated dot product Dot2 of
// Ogita, Rump and Oishi ("Accurate sum and dot product", SIAM J. Sci. Comput. 26, 2005): each product is split
// exactly into p + pe by an FMA, each running sum exactly into t + se by Knuth's TwoSum, and the error terms are
// summed beside it. The result is as accurate as a sum in 106-bit arithmetic rounded once, which is what iterative
// refinement needs to reach a working-precision solution however ill-conditioned K is. One lane per row: the call
// serves refinement steps, not the Krylov iterations.
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_residual_dot2(const int32_t *__restrict__ indptr,
                                                          const int32_t *__restrict__ indices,
                                                          const VT *__restrict__ vals, const double *__restrict__ d64,
                                                          const double *__restrict__ b,
                                                          const double *__restrict__ x, double *__restrict__ r,
                                                          int64_t n)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double s = b[i], c = 0.0;
        const int64_t d0 = i - i % 3;  // EWFE_F32_D64: first dof of the row's node (columns d0 .. d0 + 2: the block)
        for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int32_t col = indices[k];
            // row i % 3 of node i / 3's correction: 9 (i / 3) + 3 (i % 3) = 3 i; value + correction is exact
            const double a = (double)vals[k] + ((d64 && col >= d0 && col < d0 + 3) ? d64[3 * i + (col - d0)] : 0.0);
            const double y = -x[col];
            const double p = a * y;
            const double pe = fma(a, y, -p);  // a y = p + pe exactly
            const double t = s + p;
            const double z = t - s;
            const double se = (s - (t - z)) + (p - z);  // s + p = t + se exactly
            s = t;
            c += se + pe;
        }
        r[i] = s + c;
    }
}

extern "C" int ewfe_residual(const ewfe_plan *p, const void *vals, int vals_dtype, const double *b, const double *x,
                             double *r, void *stream)
{
    bool f64;
    const double *d64;
    if (!p || !p->d_indptr || !vals || !b || !x || !r || !parse_vals(p, vals, vals_dtype, f64, d64)) {
        ewfe::set_error("ewfe_residual: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t n = p->n_dof;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 65536);
    if (f64)
        hipLaunchKernelGGL(k_residual_dot2<double>, dim3(blocks), dim3(kThreads), 0, s, p->d_indptr, p->d_indices,
                           (const double *)vals, (const double *)nullptr, b, x, r, n);
    else
        hipLaunchKernelGGL(k_residual_dot2<float>, dim3(blocks), dim3(kThreads), 0, s, p->d_indptr, p->d_indices,
                           (const float *)vals, d64, b, x, r, n);
    HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int ewfe_pcg(const ewfe_plan *p, const void *vals, int vals_dtype, const double *b, double *x,
                        const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                        void *workspace, int *iters_out, double *relres_out, void *stream)
{
    bool f64;
    const double *d64;
    if (!p || !p->d_indptr || !vals || !b || !x || !workspace || maxit < 0 || check_every <= 0 ||
        !parse_vals(p, vals, vals_dtype, f64, d64) || !(rtol >= 0.0) || (n_rows > 0 && !rows_dev)) {
        ewfe::set_error("ewfe_pcg: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *ws = reinterpret_cast<double *>(workspace);
    if (f64)
        return pcg<double, false>(p, (const double *)vals, nullptr, b, x, rows_dev, n_rows, rtol, maxit, check_every,
                                  ws, iters_out, relres_out, s);
    if (d64)
        return pcg<float, true>(p, (const float *)vals, d64, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws,
                                iters_out, relres_out, s);
    return pcg<float, false>(p, (const float *)vals, nullptr, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws,
                             iters_out, relres_out, s);
}

// ------------------------------------------------------------------------------------------------
// ewfe_pcg_noalloc: ewfe_pcg with its workspace in a file-scope device array instead of a caller's buffer, for
// callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
// Every GPU holds its own copy of the array from the module load to the end of the process. In host code the
// array's name is only a stand-in for the symbol: its device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and only that address is handed to kernels. Calls on one device
// share the array, so calls that overlap on one GPU race; the scalars a solve reads before writing are cleared by
// a fill kernel at the start of every call.
// ------------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[EWFE_PCG_NOALLOC_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {nullptr};

// device address of this device's scratch array (looked up on the first call on the device)
int scratch_address(void **out)
{
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        ewfe::set_error("ewfe_pcg_noalloc: device index beyond the supported range");
        return EWFE_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) {
            ewfe::set_error("ewfe_pcg_noalloc: hipGetSymbolAddress returned no address");
            return EWFE_EHIP;
        }
        g_scratch_addr[dev] = ptr;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

__global__ void k_clear_scalars(double *scal)
{
    if (threadIdx.x < S_COUNT) scal[threadIdx.x] = 0.0;
}
}  // namespace

extern "C" int ewfe_pcg_noalloc(const ewfe_plan *p, const void *vals, int vals_dtype, const double *b, double *x,
                                const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                                int *iters_out, double *relres_out, void *stream)
{
    bool f64;
    const double *d64;
    if (!p || !p->d_indptr || !vals || !b || !x || maxit < 0 || check_every <= 0 ||
        !parse_vals(p, vals, vals_dtype, f64, d64) || !(rtol >= 0.0) || (n_rows > 0 && !rows_dev)) {
        ewfe::set_error("ewfe_pcg_noalloc: invalid arguments");
        return EWFE_EINVAL;
    }
    const int64_t need = ewfe_pcg_workspace_size(p);
    if (need > (int64_t)EWFE_PCG_NOALLOC_MAX_BYTES) {
        ewfe::set_error("ewfe_pcg_noalloc: the system needs " + std::to_string(need) +
                        " bytes of workspace, beyond the fixed scratch of " +
                        std::to_string((int64_t)EWFE_PCG_NOALLOC_MAX_BYTES) + " bytes (about " +
                        std::to_string(((int64_t)EWFE_PCG_NOALLOC_MAX_BYTES - 3 * 8 * kBlocks - 8 * S_COUNT) / 56) +
                        " dofs); use ewfe_pcg with a workspace of ewfe_pcg_workspace_size bytes");
        return EWFE_ERANGE;
    }
    void *addr = nullptr;
    int rc = scratch_address(&addr);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *ws = reinterpret_cast<double *>(addr);
    // the scalars sit after the vectors and the partial sums (the layout of pcg())
    hipLaunchKernelGGL(k_clear_scalars, dim3(1), dim3(64), 0, s, ws + 7 * p->n_dof + 3 * kBlocks);
    HIP_CHECK(hipGetLastError());
    if (f64)
        return pcg<double, false>(p, (const double *)vals, nullptr, b, x, rows_dev, n_rows, rtol, maxit, check_every,
                                  ws, iters_out, relres_out, s);
    if (d64)
        return pcg<float, true>(p, (const float *)vals, d64, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws,
                                iters_out, relres_out, s);
    return pcg<float, false>(p, (const float *)vals, nullptr, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws,
                             iters_out, relres_out, s);
}
