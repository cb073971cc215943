// libewfe.so -- CSR SpMV and a device-resident preconditioned conjugate-gradient solver for the Newton
// correction K ddU = R (the reference's NIST.linearSolve, edelweissfe/solvers/nonlinearimplicitstatic.py:727-751,
// which hands the row-eliminated matrix of applyDirichletK (:559-593) to a direct solver).
//
// The Dirichlet rows of the matrix are unit rows (ewfe_dirichlet_rows). Starting from x0 = b on those rows
// and 0 elsewhere, the initial residual vanishes on them, every search direction keeps them at zero, and
// the iteration is plain CG on the symmetric positive definite free-free block: its solution is the
// solution of the row-eliminated system. Jacobi preconditioning. All scalars stay on the device; the host
// only polls a "done" flag every few iterations. Reductions are blockwise partial sums followed by a
// single-block sum in fixed order, so the iteration history is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <string>

#include "ewfe.h"
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

constexpr int kBlocks = 1024;  // fixed grid of the row kernels -> fixed number of partial sums
constexpr int kThreads = 256;
constexpr int kLanesPerRow = 16;
constexpr int kRowsPerBlock = kThreads / kLanesPerRow;

enum Scalar { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_BB, S_DONE, S_ITERS, S_FAIL, S_COUNT = 16 };

__device__ __forceinline__ double group_sum16(double v)
{
    v += __shfl_xor(v, 8, 16);
    v += __shfl_xor(v, 4, 16);
    v += __shfl_xor(v, 2, 16);
    v += __shfl_xor(v, 1, 16);
    return v;
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

// row r of K times x, by a 16-lane group; returns the full sum on every lane of the group
template <typename VT>
__device__ __forceinline__ double row_dot(const int32_t *indptr, const int32_t *indices, const VT *vals,
                                          const double *x, int64_t r, int lane16)
{
    double acc = 0.0;
    const int32_t b = indptr[r], e = indptr[r + 1];
    for (int32_t k = b + lane16; k < e; k += kLanesPerRow) acc += (double)vals[k] * x[indices[k]];
    return group_sum16(acc);
}

template <typename VT>
__global__ __launch_bounds__(kThreads) void k_spmv(int64_t n, const int32_t *indptr, const int32_t *indices,
                                                   const VT *vals, const double *x, double *y)
{
    const int lane16 = threadIdx.x % kLanesPerRow;
    for (int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; r < n;
         r += (int64_t)gridDim.x * kRowsPerBlock) {
        const double s = row_dot(indptr, indices, vals, x, r, lane16);
        if (lane16 == 0) y[r] = s;
    }
}

// inverse of the diagonal; x0 = 0 (Dirichlet rows get b in k_set_rows)
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_pcg_setup(int64_t n, const int32_t *indptr, const int32_t *indices,
                                                        const VT *vals, double *dinv, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        double d = 1.0;
        for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k)
            if (indices[k] == r) {
                d = (double)vals[k];
                break;
            }
        dinv[r] = (d != 0.0) ? 1.0 / d : 1.0;
        x[r] = 0.0;
    }
}

__global__ void k_set_rows(const int32_t *rows, int64_t n_rows, const double *b, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride) x[rows[i]] = b[rows[i]];
}

// r = b - K x, z = dinv r, partials of (r.z, r.r, b.b)
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_pcg_residual(int64_t n, const int32_t *indptr, const int32_t *indices,
                                                           const VT *vals, const double *b, const double *x,
                                                           const double *dinv, double *r, double *z, double *p,
                                                           double *partials)
{
    const int lane16 = threadIdx.x % kLanesPerRow;
    double acc[3] = {0, 0, 0};
    for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < n;
         row += (int64_t)gridDim.x * kRowsPerBlock) {
        const double kx = row_dot(indptr, indices, vals, x, row, lane16);
        if (lane16 == 0) {
            const double bi = b[row];
            const double ri = bi - kx;
            const double zi = dinv[row] * ri;
            r[row] = ri;
            z[row] = zi;
            p[row] = zi;
            acc[0] += ri * zi;
            acc[1] += ri * ri;
            acc[2] += bi * bi;
        }
    }
    block_partials<3>(acc, partials, kBlocks);
}

// Ap = K p, partial of p.Ap
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_pcg_spmv(int64_t n, const int32_t *indptr, const int32_t *indices,
                                                       const VT *vals, const double *p, double *Ap, double *partials,
                                                       const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    const int lane16 = threadIdx.x % kLanesPerRow;
    double acc[1] = {0};
    for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanesPerRow; row < n;
         row += (int64_t)gridDim.x * kRowsPerBlock) {
        const double s = row_dot(indptr, indices, vals, p, row, lane16);
        if (lane16 == 0) {
            Ap[row] = s;
            acc[0] += s * p[row];
        }
    }
    block_partials<1>(acc, partials, kBlocks);
}

// x += alpha p, r -= alpha Ap, z = dinv r; partials of (r.z, r.r)
__global__ __launch_bounds__(kThreads) void k_pcg_update(int64_t n, double *x, double *r, double *z, const double *p,
                                                         const double *Ap, const double *dinv, double *partials,
                                                         const double *scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double alpha = scal[S_ALPHA];
    double acc[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        const double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        acc[0] += ri * zi;
        acc[1] += ri * ri;
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

template <typename VT>
int pcg(const ewfe_plan *p, const VT *vals, const double *b, double *x, const int32_t *rows, int64_t n_rows,
        double rtol, int maxit, int check_every, double *ws, int *iters_out, double *relres_out, hipStream_t s)
{
    const int64_t n = p->n_dof;
    double *r = ws, *z = ws + n, *pp = ws + 2 * n, *Ap = ws + 3 * n, *dinv = ws + 4 * n;
    double *partials = ws + 5 * n, *scal = partials + 3 * kBlocks;
    const double rtol2 = rtol * rtol;
    hipLaunchKernelGGL(k_pcg_setup<VT>, dim3(kBlocks), dim3(kThreads), 0, s, n, p->d_indptr, p->d_indices, vals,
                       dinv, x);
    if (n_rows > 0)
        hipLaunchKernelGGL(k_set_rows, dim3(256), dim3(256), 0, s, rows, n_rows, b, x);
    hipLaunchKernelGGL(k_pcg_residual<VT>, dim3(kBlocks), dim3(kThreads), 0, s, n, p->d_indptr, p->d_indices, vals,
                       b, (const double *)x, (const double *)dinv, r, z, pp, partials);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 0, (const double *)partials, scal, rtol2, maxit);
    HIP_CHECK(hipGetLastError());
    double host_scal[S_COUNT];
    int it = 0;
    for (;;) {
        HIP_CHECK(hipMemcpyAsync(host_scal, scal, sizeof(host_scal), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (host_scal[S_DONE] != 0.0 || it >= maxit) break;
        for (int k = 0; k < check_every && it < maxit; ++k, ++it) {
            hipLaunchKernelGGL(k_pcg_spmv<VT>, dim3(kBlocks), dim3(kThreads), 0, s, n, p->d_indptr, p->d_indices,
                               vals, (const double *)pp, Ap, partials, (const double *)scal);
            hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(1024), 0, s, 1, (const double *)partials, scal, rtol2,
                               maxit);
            hipLaunchKernelGGL(k_pcg_update, dim3(kBlocks), dim3(kThreads), 0, s, n, x, r, z, (const double *)pp,
                               (const double *)Ap, (const double *)dinv, partials, (const double *)scal);
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

}  // namespace

extern "C" int64_t ewfe_pcg_workspace_size(const ewfe_plan *p)
{
    if (!p) return 0;
    return (int64_t)sizeof(double) * (5 * p->n_dof + 3 * kBlocks + S_COUNT);
}

extern "C" int ewfe_spmv(const ewfe_plan *p, const void *vals, int vals_dtype, const double *x, double *y,
                         void *stream)
{
    if (!p || !p->d_indptr || !vals || !x || !y || (vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32)) {
        ewfe::set_error("ewfe_spmv: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (vals_dtype == EWFE_F64)
        hipLaunchKernelGGL(k_spmv<double>, dim3(kBlocks), dim3(kThreads), 0, s, p->n_dof, p->d_indptr, p->d_indices,
                           (const double *)vals, x, y);
    else
        hipLaunchKernelGGL(k_spmv<float>, dim3(kBlocks), dim3(kThreads), 0, s, p->n_dof, p->d_indptr, p->d_indices,
                           (const float *)vals, x, y);
    HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int ewfe_pcg(const ewfe_plan *p, const void *vals, int vals_dtype, const double *b, double *x,
                        const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                        void *workspace, int *iters_out, double *relres_out, void *stream)
{
    if (!p || !p->d_indptr || !vals || !b || !x || !workspace || maxit < 0 || check_every <= 0 ||
        (vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32) || !(rtol >= 0.0) || (n_rows > 0 && !rows_dev)) {
        ewfe::set_error("ewfe_pcg: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *ws = reinterpret_cast<double *>(workspace);
    if (vals_dtype == EWFE_F64)
        return pcg<double>(p, (const double *)vals, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws, iters_out,
                           relres_out, s);
    return pcg<float>(p, (const float *)vals, b, x, rows_dev, n_rows, rtol, maxit, check_every, ws, iters_out,
                      relres_out, s);
}
