// libewfe.so -- smoothed-aggregation multigrid (SA-AMG) preconditioner for the Newton correction K ddU = R.
//
// The reference's iterative option is GMRES preconditioned by pyamg's smoothed-aggregation solver
// (edelweissfe/linsolve/gmres/gmres.py:70-88, registered in edelweissfe/config/linsolve.py:46-86). Here the same
// method runs on the device as a symmetric V-cycle inside CG (ewfe_pcg_mg in ewfe_solver.hip):
//   * unknowns grouped by node (3x3 blocks); near-null space = the three translations, so the tentative
//     prolongator P0 is a block identity per aggregate;
//   * P = (I - w D^-1 A) P0, w = 4 / (3 lambda_max(D^-1 A)), D = 3x3 block diagonal, lambda_max by power
//     iteration on the device;
//   * A_c = P^T (A P), two sparse products with patterns built once per plan on the host
//     (ewfe_amg_host.cpp), each output block summed by one lane in a fixed order (bitwise reproducible);
//   * smoother: Chebyshev polynomial of degree nu in D^-1 A on [lambda_max / ratio, 1.1 lambda_max]
//     (Adams, Brezina, Hu, Tuminaro 2003), the same polynomial before and after the coarse correction, so the
//     V-cycle is a symmetric positive definite preconditioner;
//   * coarsest level (<= a few hundred nodes): dense inverse, computed on the host by Cholesky;
//   * the prolongators are stored in the precision of the fine matrix (FP32 for an FP32 tangent: the
//     preconditioner stays a fixed symmetric positive definite operator, at half the transfer traffic);
//     Galerkin products, coarse operators, smoother inverses and all vectors are FP64.
// Dirichlet dofs: the fine matrix has unit rows there (row elimination). The prolongator's rows of Dirichlet
// dofs are zero (P0 masked, and the masked columns of A dropped), so every level works on the free dofs only;
// the smoother keeps vectors at zero on Dirichlet dofs because D^-1 has identity rows there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

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

constexpr int kBlocks = 4096;
constexpr int kThreads = 256;
constexpr int kGroup = 32;  // lanes per node in the SpMV-type kernels
constexpr int kNodesPerBlock = kThreads / kGroup;
constexpr int kPowerIters = 15;
using ewfe::kAmgMaxP;
using ewfe::kAmgMaxRow;

struct MgLevel {
    int64_t n = 0;  // nodes
    // matrix of levels >= 1 (block CSR, row-major 3x3 blocks)
    int32_t *aptr = nullptr, *acol = nullptr, *adiag = nullptr;
    double *aval = nullptr;
    int64_t nnzb = 0;
    double *dinv = nullptr;  // [n][9]
    // transfer to the next level
    int64_t n_agg = 0;
    int32_t *pptr = nullptr, *pcol = nullptr;
    uint8_t *pslot = nullptr;
    double *pval = nullptr;
    int32_t *wptr = nullptr, *wcol = nullptr;
    double *wval = nullptr;
    int32_t *rptr = nullptr, *ri = nullptr, *rt = nullptr;
    double *rval = nullptr;  // P's blocks in restriction order, rval[q] = pval[rt[q]] (set up with P; FP32 P: floats)
    int64_t n_r = 0;         // entries of the restriction lists
    // V-cycle vectors (3n each); level 0 uses the caller's right-hand side / result for b / x
    double *b = nullptr, *x = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr;
};

}  // namespace

struct ewfe_mg {
    const ewfe_plan *plan = nullptr;
    int device = 0;
    std::vector<MgLevel> lv;
    // coarsest level: host pattern for the dense assembly, device dense inverse
    std::vector<int32_t> c_aptr, c_acol;
    double *ainv = nullptr;
    int64_t nc3 = 0;
    double *lam = nullptr;       // [levels] lambda_max(D^-1 A)
    double *partials = nullptr;  // [kBlocks]
    uint8_t *mask = nullptr;     // [n_dof] 1 on Dirichlet dofs
    double *pw_x = nullptr, *pw_y = nullptr;
    int degree = 2;
    double ratio = 20.0;
    bool ready = false;
    int setup_dtype = -1;  // precision of the matrix the hierarchy was set up for (= prolongator storage)
    // fine-level passes of the V-cycle in half precision (ewfe_mg_set_fine_precision): the set-up matrix divided by
    // a power of two (its largest magnitude then lies in [2^14, 2^15)) and rounded to FP16
    bool fine16 = false;
    _Float16 *vals16 = nullptr;
    double *scale16 = nullptr;  // [2]: the largest magnitude, the scale
    // FP32 matrices: the fine level's block-Jacobi inverses rounded to FP32 for the V-cycle's smoothing steps
    // (allocated by the first FP32 setup)
    float *dinv32 = nullptr;
    std::vector<void *> allocs;
};

namespace {

__device__ __forceinline__ double group_sum32(double v)
{
    v += __shfl_xor(v, 16, 32);
    v += __shfl_xor(v, 8, 32);
    v += __shfl_xor(v, 4, 32);
    v += __shfl_xor(v, 2, 32);
    v += __shfl_xor(v, 1, 32);
    return v;
}

__device__ __forceinline__ double wave_sum64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ void block_partial1(double v, double *partials)
{
    __shared__ double red[kThreads / 64];
    v = wave_sum64(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kThreads / 64; ++w) s += red[w];
        partials[blockIdx.x] = s;
    }
}

// ---- level operators --------------------------------------------------------------------------------
// Fine level: the plan's CSR; row node i owns dofs 3i..3i+2; rn_meta = {first dof, CSR offset, nb_ptr, degree};
// nbcol = first dof of each neighbour block. Coarse levels: block CSR.
template <typename VT>
struct FineOp {
    const int4 *meta;
    const int32_t *nbcol, *diag_rank;
    const VT *vals;
    const uint8_t *mask;
    int64_t n;
    const double *scale = nullptr;  // half-precision values: [1] is the scale they were divided by
};

struct BsrOp {
    const int32_t *aptr, *acol, *adiag;
    const double *aval;
    int64_t n;
};

template <typename VT, typename XT>
__device__ __forceinline__ void row_apply(const FineOp<VT> &A, int64_t i, int lane, const XT *__restrict__ x,
                                          double &y0, double &y1, double &y2)
{
    const int4 m4 = A.meta[i];
    const int32_t base = m4.y, nb0 = m4.z, deg = m4.w, L = 3 * deg;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int m = lane; m < deg; m += kGroup) {
        const int32_t c = A.nbcol[nb0 + m];
        const double x0 = (double)x[c], x1 = (double)x[c + 1], x2 = (double)x[c + 2];
        const VT *v0 = A.vals + base + 3 * m;
        a0 += (double)v0[0] * x0 + (double)v0[1] * x1 + (double)v0[2] * x2;
        a1 += (double)v0[L] * x0 + (double)v0[L + 1] * x1 + (double)v0[L + 2] * x2;
        a2 += (double)v0[2 * L] * x0 + (double)v0[2 * L + 1] * x1 + (double)v0[2 * L + 2] * x2;
    }
    y0 = group_sum32(a0);
    y1 = group_sum32(a1);
    y2 = group_sum32(a2);
    if constexpr (std::is_same<VT, _Float16>::value) {
        const double s = A.scale[1];
        y0 *= s;
        y1 *= s;
        y2 *= s;
    }
}

template <typename XT>
__device__ __forceinline__ void row_apply(const BsrOp &A, int64_t i, int lane, const XT *__restrict__ x,
                                          double &y0, double &y1, double &y2)
{
    double a0 = 0, a1 = 0, a2 = 0;
    for (int32_t e = A.aptr[i] + lane; e < A.aptr[i + 1]; e += kGroup) {
        const int32_t c = 3 * A.acol[e];
        const double x0 = (double)x[c], x1 = (double)x[c + 1], x2 = (double)x[c + 2];
        const double *v = A.aval + 9 * (int64_t)e;
        a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
        a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
        a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    y0 = group_sum32(a0);
    y1 = group_sum32(a1);
    y2 = group_sum32(a2);
}

template <typename DT>
__device__ __forceinline__ void apply3(const DT *__restrict__ Q, double r0, double r1, double r2, double &z0,
                                       double &z1, double &z2)
{
    z0 = (double)Q[0] * r0 + (double)Q[1] * r1 + (double)Q[2] * r2;
    z1 = (double)Q[3] * r0 + (double)Q[4] * r1 + (double)Q[5] * r2;
    z2 = (double)Q[6] * r0 + (double)Q[7] * r1 + (double)Q[8] * r2;
}

__global__ void k_mg_narrow(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (float)a[i];
}

// inverse of a 3x3 diagonal block; a component whose row and column vanish (a coarse dof that only Dirichlet
// dofs feed) gets a unit diagonal, a singular remainder falls back to point Jacobi
__device__ void block_inverse(double B[3][3], double *Q)
{
    for (int c = 0; c < 3; ++c) {
        bool zero = true;
        for (int t = 0; t < 3; ++t) zero = zero && B[c][t] == 0.0 && B[t][c] == 0.0;
        if (zero) B[c][c] = 1.0;
    }
    const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                       B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                       B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
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
    } else {
        for (int t = 0; t < 9; ++t) Q[t] = 0.0;
        for (int c = 0; c < 3; ++c) Q[4 * c] = B[c][c] != 0.0 ? 1.0 / B[c][c] : 1.0;
    }
}

// ---- setup kernels ----------------------------------------------------------------------------------
__global__ void k_mg_mask(const int32_t *rows, int64_t n_rows, uint8_t *mask)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride) mask[rows[i]] = 1;
}

// ---- half-precision copy of the fine-level matrix ---------------------------------------------------------------
// scale16[0] = max |v| (positive doubles order like their bit patterns, so the maximum is an integer atomic max and
// does not depend on the order of the updates)
template <typename VT>
__global__ __launch_bounds__(256) void k_mg_absmax(const VT *__restrict__ v, int64_t n, double *scale16)
{
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        m = fmax(m, fabs((double)v[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(scale16), (unsigned long long)__double_as_longlong(m));
}

// the scale: a power of two with max |v| / scale in [2^14, 2^15) (well inside FP16's range, 65504)
__device__ __forceinline__ double half_scale(double vmax)
{
    if (!(vmax > 0.0)) return 1.0;
    int e;
    (void)frexp(vmax, &e);  // vmax = f 2^e, f in [0.5, 1)
    return ldexp(1.0, e - 15);
}

template <typename VT>
__global__ __launch_bounds__(256) void k_mg_to_half(const VT *__restrict__ v, int64_t n, double *scale16,
                                                    _Float16 *__restrict__ h)
{
    const double s = half_scale(scale16[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) scale16[1] = s;
    const double inv = 1.0 / s;  // exact: s is a power of two
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        h[i] = (_Float16)((double)v[i] * inv);
}

__global__ void k_mg_zero2(double *p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0.0;
}

template <typename VT>
__global__ void k_mg_dinv_fine(FineOp<VT> A, double *dinv)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {
        const int4 m4 = A.meta[i];
        const int32_t L = 3 * m4.w, off = m4.y + 3 * A.diag_rank[i];
        double B[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[r][c] = (double)A.vals[off + r * L + c];
        block_inverse(B, dinv + 9 * i);
    }
}

__global__ void k_mg_dinv_bsr(BsrOp A, double *dinv)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {
        const double *v = A.aval + 9 * (int64_t)A.adiag[i];
        double B[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[r][c] = v[3 * r + c];
        block_inverse(B, dinv + 9 * i);
    }
}

// power iteration for lambda_max(D^-1 A): start vector in [0.5, 1.5) (a fixed hash), zero on Dirichlet dofs
__global__ void k_mg_power_init(int64_t n3, const uint8_t *mask, double *x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += stride) {
        uint64_t h = (uint64_t)k * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 32;
        const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        x[k] = (mask && mask[k]) ? 0.0 : 0.5 + u;
    }
}

template <typename OP>
__global__ __launch_bounds__(kThreads) void k_mg_power(OP A, const double *dinv, const double *x, double *y,
                                                       double *partials)
{
    const int lane = threadIdx.x % kGroup;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kGroup; i < A.n;
         i += (int64_t)gridDim.x * kNodesPerBlock) {
        double a0, a1, a2;
        row_apply(A, i, lane, x, a0, a1, a2);  // every lane of the group holds the sums
        if (lane < 3) {                        // lane c: component c
            const double *Q = dinv + 9 * i + 3 * lane;
            const double z = Q[0] * a0 + Q[1] * a1 + Q[2] * a2;
            y[3 * i + lane] = z;
            acc += z * z;
        }
    }
    block_partial1(acc, partials);
}

__global__ __launch_bounds__(1024) void k_mg_norm(const double *partials, double *out)
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
    if (threadIdx.x == 0) *out = sqrt(red[0]);
}

__global__ void k_mg_scale(int64_t n3, const double *y, double *x, const double *nrm)
{
    const double s = *nrm > 0.0 ? 1.0 / *nrm : 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += stride) x[k] = y[k] * s;
}

// P = (I - w D^-1 A) P0 for row node i (one 64-lane wave per node). Fine level: columns of Dirichlet dofs are
// dropped from A and the Dirichlet rows of P0 are zero, so P vanishes on Dirichlet rows.
template <typename VT>
__device__ __forceinline__ void load_block_fine(const FineOp<VT> &A, int32_t base, int32_t L, int m, int32_t col,
                                                double *dst)
{
    const VT *v = A.vals + base + 3 * m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) dst[3 * r + c] = A.mask[col + c] ? 0.0 : (double)v[r * L + c];
}

template <bool FINE, typename VT, typename PT>
__global__ __launch_bounds__(64) void k_mg_prolongator(FineOp<VT> F, BsrOp B, const double *__restrict__ dinv,
                                                       const int32_t *__restrict__ pptr,
                                                       const uint8_t *__restrict__ pslot, PT *__restrict__ pval,
                                                       const double *__restrict__ lam)
{
    __shared__ double s_a[kAmgMaxRow * 9];
    __shared__ uint8_t s_slot[kAmgMaxRow];
    __shared__ double s_s[kAmgMaxP * 9];
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x;
    int32_t deg, e0, self;
    if (FINE) {
        const int4 m4 = F.meta[i];
        deg = m4.w;
        e0 = m4.z;
        self = F.diag_rank[i];
        for (int m = lane; m < deg; m += 64) load_block_fine(F, m4.y, 3 * deg, m, F.nbcol[e0 + m], s_a + 9 * m);
    } else {
        e0 = B.aptr[i];
        deg = B.aptr[i + 1] - e0;
        self = B.adiag[i] - e0;
        for (int t = lane; t < 9 * deg; t += 64) s_a[t] = B.aval[9 * (int64_t)e0 + t];
    }
    for (int m = lane; m < deg; m += 64) s_slot[m] = pslot[e0 + m];
    const int32_t p0 = pptr[i], np = pptr[i + 1] - p0;
    __syncthreads();
    // S_t = sum of the (masked) blocks whose column node lies in aggregate t, ascending neighbour order
    for (int q = lane; q < 9 * np; q += 64) {
        const int t = q / 9, rc = q % 9;
        double s = 0.0;
        for (int m = 0; m < deg; ++m)
            if (s_slot[m] == t) s += s_a[9 * m + rc];
        s_s[q] = s;
    }
    __syncthreads();
    const double w = 4.0 / (3.0 * lam[0]);
    const int tself = s_slot[self];
    const double *Q = dinv + 9 * i;
    for (int q = lane; q < 9 * np; q += 64) {
        const int t = q / 9, r = (q % 9) / 3, c = q % 3;
        double v = -w * (Q[3 * r] * s_s[9 * t + c] + Q[3 * r + 1] * s_s[9 * t + 3 + c] + Q[3 * r + 2] * s_s[9 * t + 6 + c]);
        if (t == tself && r == c && !(FINE && F.mask[3 * i + r])) v += 1.0;
        pval[9 * (int64_t)(p0 + t) + 3 * r + c] = (PT)v;
    }
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int32_t wave_scan64(int32_t v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// W = A P, row node i (one 64-lane wave): lane u owns output block u and sums A_ik P_kJ over the neighbours k in
// ascending order. The row of A and the neighbours' P column indices are staged in LDS; the P blocks are staged
// too when STAGE_PV (coarse levels), else read from L2 for the matching entries only (fine level: the small LDS
// footprint keeps many workgroups per CU in flight). Rows whose neighbours' P rows exceed MAXSTAGE read P from
// L2 altogether. The grid walks XCD-contiguous node ranges (workgroups are dealt to the 8 XCDs round robin).
template <bool FINE, typename VT, typename PT, int MAXROW, int MAXSTAGE, bool STAGE_PV>
__global__ __launch_bounds__(64) void k_mg_ap(FineOp<VT> F, BsrOp B, const int32_t *__restrict__ pptr,
                                              const int32_t *__restrict__ pcol, const PT *__restrict__ pval,
                                              const int32_t *__restrict__ wptr, const int32_t *__restrict__ wcol,
                                              double *__restrict__ wval, int64_t n, int64_t xcd_chunk)
{
    __shared__ double s_a[MAXROW * 9];
    __shared__ int32_t s_k[MAXROW];
    __shared__ int32_t s_off[MAXROW + 1];
    __shared__ int32_t s_pc[MAXSTAGE];
    __shared__ int32_t s_t[MAXSTAGE];
    __shared__ double s_pv[STAGE_PV ? MAXSTAGE * 9 : 1];
    const int64_t i = (int64_t)(blockIdx.x & 7) * xcd_chunk + (blockIdx.x >> 3);
    if (i >= n) return;  // uniform over the workgroup
    const int lane = threadIdx.x;
    int32_t deg;
    if (FINE) {
        const int4 m4 = F.meta[i];
        deg = m4.w;
        const int32_t L = 3 * deg;
        for (int m = lane; m < deg; m += 64) {
            const VT *v = F.vals + m4.y + 3 * m;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) s_a[9 * m + 3 * r + c] = (double)v[r * L + c];
            s_k[m] = F.nbcol[m4.z + m] / 3;
        }
    } else {
        const int32_t e0 = B.aptr[i];
        deg = B.aptr[i + 1] - e0;
        for (int t = lane; t < 9 * deg; t += 64) s_a[t] = B.aval[9 * (int64_t)e0 + t];
        for (int m = lane; m < deg; m += 64) s_k[m] = B.acol[e0 + m];
    }
    // neighbour m's P row length, prefix-summed (rows of up to MAXROW <= 128 neighbours: two wave passes)
    int32_t carry = 0;
    if (lane == 0) s_off[0] = 0;
    for (int m0 = 0; m0 < deg; m0 += 64) {
        const int m = m0 + lane;
        int32_t cnt = 0;
        if (m < deg) {
            const int32_t k = FINE ? F.nbcol[F.meta[i].z + m] / 3 : B.acol[B.aptr[i] + m];
            cnt = pptr[k + 1] - pptr[k];
        }
        const int32_t incl = wave_scan64(cnt) + carry;
        if (m < deg) s_off[m + 1] = incl;
        carry = __shfl(incl, 63, 64);
    }
    __syncthreads();
    const int32_t total = s_off[deg];
    const int32_t w0 = wptr[i], nw = wptr[i + 1] - w0;
    if (total > MAXSTAGE) {  // rows beyond the staging capacity read the neighbours' P rows from L2
        for (int u = lane; u < nw; u += 64) {
            const int32_t J = wcol[w0 + u];
            double acc[9];
            for (int t = 0; t < 9; ++t) acc[t] = 0.0;
            for (int m = 0; m < deg; ++m) {
                const int32_t k = s_k[m];
                for (int32_t t = pptr[k]; t < pptr[k + 1]; ++t) {
                    const int32_t c = pcol[t];
                    if (c < J) continue;
                    if (c == J) {
                        const double *a = s_a + 9 * m;
                        const PT *p = pval + 9 * (int64_t)t;
                        for (int r = 0; r < 3; ++r)
                            for (int cc = 0; cc < 3; ++cc)
                                acc[3 * r + cc] += a[3 * r] * (double)p[cc] + a[3 * r + 1] * (double)p[3 + cc] +
                                                   a[3 * r + 2] * (double)p[6 + cc];
                    }
                    break;
                }
            }
            for (int t = 0; t < 9; ++t) wval[9 * (int64_t)(w0 + u) + t] = acc[t];
        }
        return;  // uniform over the workgroup
    }
    for (int32_t g = lane; g < total; g += 64) {  // stage the neighbours' P rows, one block per lane
        int lo = 0, hi = deg;                      // m with s_off[m] <= g < s_off[m + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= g) lo = mid;
            else hi = mid;
        }
        const int32_t t = pptr[s_k[lo]] + (g - s_off[lo]);
        s_pc[g] = pcol[t];
        s_t[g] = t;
        if (STAGE_PV)
            for (int q = 0; q < 9; ++q) s_pv[9 * g + q] = (double)pval[9 * (int64_t)t + q];
    }
    __syncthreads();
    for (int u = lane; u < nw; u += 64) {
        const int32_t J = wcol[w0 + u];
        double acc[9];
        for (int t = 0; t < 9; ++t) acc[t] = 0.0;
        for (int m = 0; m < deg; ++m) {
            for (int32_t o = s_off[m]; o < s_off[m + 1]; ++o) {
                const int32_t c = s_pc[o];
                if (c < J) continue;
                if (c == J) {
                    const double *a = s_a + 9 * m;
                    double p[9];
                    for (int q = 0; q < 9; ++q) p[q] = STAGE_PV ? s_pv[9 * o + q] : (double)pval[9 * (int64_t)s_t[o] + q];
                    for (int r = 0; r < 3; ++r)
                        for (int cc = 0; cc < 3; ++cc)
                            acc[3 * r + cc] += a[3 * r] * p[cc] + a[3 * r + 1] * p[3 + cc] + a[3 * r + 2] * p[6 + cc];
                }
                break;  // P rows are sorted
            }
        }
        for (int t = 0; t < 9; ++t) wval[9 * (int64_t)(w0 + u) + t] = acc[t];
    }
}
constexpr int kFineMaxRow = 64, kFineMaxStage = 256;      // level 0, 8-node meshes: LDS ~7 KB per workgroup
constexpr int kWideMaxRow = 128, kWideMaxStage = 512;      // level 0, 20-node meshes (up to 81 blocks): ~15 KB
constexpr int kCoarseMaxRow = 128, kCoarseMaxStage = 512;  // levels >= 1: LDS ~51 KB

// A_c = P^T W, aggregate I (one 64-lane wave): lane v owns output block v (column J) and sums P_tI^T W_iJ over
// the rows i of R(I) in ascending order; W_iJ is found by binary search in W's sorted row.
template <typename PT>
__global__ __launch_bounds__(64) void k_mg_rap(int64_t n_agg, const int32_t *__restrict__ rptr,
                                               const int32_t *__restrict__ ri, const int32_t *__restrict__ rt,
                                               const PT *__restrict__ pval, const int32_t *__restrict__ wptr,
                                               const int32_t *__restrict__ wcol, const double *__restrict__ wval,
                                               const int32_t *__restrict__ captr, const int32_t *__restrict__ cacol,
                                               double *__restrict__ caval)
{
    const int64_t I = blockIdx.x;
    const int lane = threadIdx.x;
    const int32_t c0 = captr[I], nc = captr[I + 1] - c0;
    const int32_t q0 = rptr[I], q1 = rptr[I + 1];
    for (int v = lane; v < nc; v += 64) {
        const int32_t J = cacol[c0 + v];
        double acc[9];
        for (int t = 0; t < 9; ++t) acc[t] = 0.0;
        for (int32_t q = q0; q < q1; ++q) {
            const int32_t i = ri[q];
            int32_t lo = wptr[i], hi = wptr[i + 1];
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (wcol[mid] < J) lo = mid + 1;
                else hi = mid;
            }
            if (lo == wptr[i + 1] || wcol[lo] != J) continue;
            const PT *p = pval + 9 * (int64_t)rt[q];
            const double *w = wval + 9 * (int64_t)lo;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    acc[3 * r + c] += (double)p[r] * w[c] + (double)p[3 + r] * w[3 + c] + (double)p[6 + r] * w[6 + c];
        }
        for (int t = 0; t < 9; ++t) caval[9 * (int64_t)(c0 + v) + t] = acc[t];
    }
}

// ---- V-cycle kernels ------------------------------------------------------------------------------------
struct Cheb {
    double theta, delta, sigma;
};

__device__ __forceinline__ Cheb cheb_params(double lam, double ratio)
{
    const double hi = 1.1 * lam, lo = hi / ratio;
    Cheb c;
    c.theta = 0.5 * (hi + lo);
    c.delta = 0.5 * (hi - lo);
    c.sigma = c.theta / c.delta;
    return c;
}

// rho_k of the Chebyshev recurrence: rho_0 = 1 / sigma, rho_k = 1 / (2 sigma - rho_{k-1})
__device__ __forceinline__ double cheb_rho(const Cheb &c, int k)
{
    double rho = 1.0 / c.sigma;
    for (int j = 1; j <= k; ++j) rho = 1.0 / (2.0 * c.sigma - rho);
    return rho;
}

// first step from x = 0: r = b, d = D^-1 b / theta, x = d (vectors of type XT: double, or float on the fine level of
// an FP32 matrix)
template <typename DT, typename XT>
__global__ void k_mg_cheb0_zero(int64_t n, const DT *dinv, const double *b, XT *r, XT *d, XT *x,
                                const double *lam, double ratio, const double *done)
{
    if (done && *done != 0.0) return;
    const Cheb c = cheb_params(*lam, ratio);
    const double s = 1.0 / c.theta;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double b0 = b[3 * i], b1 = b[3 * i + 1], b2 = b[3 * i + 2];
        double z0, z1, z2;
        apply3(dinv + 9 * i, b0, b1, b2, z0, z1, z2);
        r[3 * i] = (XT)b0, r[3 * i + 1] = (XT)b1, r[3 * i + 2] = (XT)b2;
        d[3 * i] = (XT)(s * z0), d[3 * i + 1] = (XT)(s * z1), d[3 * i + 2] = (XT)(s * z2);
        x[3 * i] = (XT)(s * z0), x[3 * i + 1] = (XT)(s * z1), x[3 * i + 2] = (XT)(s * z2);
    }
}

// first step from a given x (r = b - A x already computed): d = D^-1 r / theta, x += d
template <typename DT, typename XT>
__global__ void k_mg_cheb0(int64_t n, const DT *dinv, const XT *r, XT *d, XT *x, const double *lam,
                           double ratio, const double *done)
{
    if (done && *done != 0.0) return;
    const Cheb c = cheb_params(*lam, ratio);
    const double s = 1.0 / c.theta;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double z0, z1, z2;
        apply3(dinv + 9 * i, (double)r[3 * i], (double)r[3 * i + 1], (double)r[3 * i + 2], z0, z1, z2);
        d[3 * i] = (XT)(s * z0), d[3 * i + 1] = (XT)(s * z1), d[3 * i + 2] = (XT)(s * z2);
        x[3 * i] = (XT)((double)x[3 * i] + s * z0), x[3 * i + 1] = (XT)((double)x[3 * i + 1] + s * z1),
        x[3 * i + 2] = (XT)((double)x[3 * i + 2] + s * z2);
    }
}

// r = b - A x (b may alias r: each node reads only its own entries of b)
template <typename OP, typename BT, typename XT>
__global__ __launch_bounds__(kThreads) void k_mg_resid(OP A, const BT *b, const XT *__restrict__ x, XT *r,
                                                       const double *done)
{
    if (done && *done != 0.0) return;
    const int lane = threadIdx.x % kGroup;
    for (int64_t i = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kGroup; i < A.n;
         i += (int64_t)gridDim.x * kNodesPerBlock) {
        double y0, y1, y2;
        row_apply(A, i, lane, x, y0, y1, y2);
        if (lane < 3) r[3 * i + lane] = (XT)((double)b[3 * i + lane] - (lane == 0 ? y0 : (lane == 1 ? y1 : y2)));
    }
}

// Chebyshev step k >= 1: r -= A d_old, d_new = rho_k rho_{k-1} d_old + (2 rho_k / delta) D^-1 r, x += d_new
template <typename OP, typename DT, typename XT>
__global__ __launch_bounds__(kThreads) void k_mg_cheb_step(OP A, const DT *dinv, XT *r, const XT *__restrict__ d_old,
                                                           XT *d_new, XT *x, const double *lam, double ratio, int k,
                                                           const double *done)
{
    if (done && *done != 0.0) return;
    const Cheb c = cheb_params(*lam, ratio);
    const double rho_prev = cheb_rho(c, k - 1), rho = cheb_rho(c, k);
    const double ca = rho * rho_prev, cb = 2.0 * rho / c.delta;
    const int lane = threadIdx.x % kGroup;
    for (int64_t i = (int64_t)blockIdx.x * kNodesPerBlock + threadIdx.x / kGroup; i < A.n;
         i += (int64_t)gridDim.x * kNodesPerBlock) {
        double y0, y1, y2;
        row_apply(A, i, lane, d_old, y0, y1, y2);
        // lanes 0..2 each own one component; the 3x3 block-Jacobi product takes the three residual
        // components from those lanes by shuffles (the other lanes carry copies of component 2)
        const int c = lane < 3 ? lane : 2;
        const double rc = (double)r[3 * i + c] - (c == 0 ? y0 : (c == 1 ? y1 : y2));
        const double r0 = __shfl(rc, 0, kGroup), r1 = __shfl(rc, 1, kGroup), r2 = __shfl(rc, 2, kGroup);
        if (lane < 3) {
            const DT *Q = dinv + 9 * i + 3 * c;
            const double e =
                ca * (double)d_old[3 * i + c] + cb * ((double)Q[0] * r0 + (double)Q[1] * r1 + (double)Q[2] * r2);
            r[3 * i + c] = (XT)rc;
            d_new[3 * i + c] = (XT)e;
            x[3 * i + c] = (XT)((double)x[3 * i + c] + e);
        }
    }
}

// rval[q] = pval[rt[q]]: P's blocks in the order restriction reads them (contiguous per aggregate)
template <typename PT>
__global__ void k_mg_rgather(int64_t n_r, const int32_t *__restrict__ rt, const PT *__restrict__ pval,
                             PT *__restrict__ rval)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < 9 * n_r; k += stride)
        rval[k] = pval[9 * (int64_t)rt[k / 9] + k % 9];
}

// b_c[I] = sum over R(I) of P_t^T r[i] (one 64-lane wave per aggregate, fixed-order butterfly); the blocks from the
// restriction-ordered copy rval
template <typename PT, typename XT>
__global__ __launch_bounds__(64) void k_mg_restrict(int64_t n_agg, const int32_t *__restrict__ rptr,
                                                    const int32_t *__restrict__ ri, const PT *__restrict__ rval,
                                                    const XT *__restrict__ r, double *__restrict__ bc,
                                                    const double *done)
{
    if (done && *done != 0.0) return;
    const int64_t I = blockIdx.x;
    const int lane = threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int32_t q = rptr[I] + lane; q < rptr[I + 1]; q += 64) {
        const int32_t i = ri[q];
        const PT *p = rval + 9 * (int64_t)q;
        const double r0 = (double)r[3 * i], r1 = (double)r[3 * i + 1], r2 = (double)r[3 * i + 2];
        a0 += (double)p[0] * r0 + (double)p[3] * r1 + (double)p[6] * r2;
        a1 += (double)p[1] * r0 + (double)p[4] * r1 + (double)p[7] * r2;
        a2 += (double)p[2] * r0 + (double)p[5] * r1 + (double)p[8] * r2;
    }
    a0 = wave_sum64(a0);
    a1 = wave_sum64(a1);
    a2 = wave_sum64(a2);
    if (lane == 0) bc[3 * I] = a0, bc[3 * I + 1] = a1, bc[3 * I + 2] = a2;
}

// x[i] += sum_t P_t x_c[J_t]
template <typename PT, typename XT>
__global__ void k_mg_prolong(int64_t n, const int32_t *__restrict__ pptr, const int32_t *__restrict__ pcol,
                             const PT *__restrict__ pval, const double *__restrict__ xc, XT *__restrict__ x,
                             const double *done)
{
    if (done && *done != 0.0) return;
    // 4 lanes per fine node: lane s takes blocks s, s + 4, ... of the node's row (neighbouring lanes read neighbouring
    // blocks), a fixed 2-step butterfly sums them (blockDim is a multiple of 4, so a node's lanes share a wave)
    const int sub = threadIdx.x & 3;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 2;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2; i < n; i += stride) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int32_t t = pptr[i] + sub; t < pptr[i + 1]; t += 4) {
            const int32_t J = 3 * pcol[t];
            const PT *p = pval + 9 * (int64_t)t;
            const double c0 = xc[J], c1 = xc[J + 1], c2 = xc[J + 2];
            a0 += (double)p[0] * c0 + (double)p[1] * c1 + (double)p[2] * c2;
            a1 += (double)p[3] * c0 + (double)p[4] * c1 + (double)p[5] * c2;
            a2 += (double)p[6] * c0 + (double)p[7] * c1 + (double)p[8] * c2;
        }
        a0 += __shfl_xor(a0, 1, 4), a1 += __shfl_xor(a1, 1, 4), a2 += __shfl_xor(a2, 1, 4);
        a0 += __shfl_xor(a0, 2, 4), a1 += __shfl_xor(a1, 2, 4), a2 += __shfl_xor(a2, 2, 4);
        if (sub == 0)
            x[3 * i] = (XT)((double)x[3 * i] + a0), x[3 * i + 1] = (XT)((double)x[3 * i + 1] + a1),
            x[3 * i + 2] = (XT)((double)x[3 * i + 2] + a2);
    }
}

// z = x (the fine level's FP32 iterate handed back in FP64)
__global__ void k_mg_widen(int64_t n3, const float *__restrict__ x, double *__restrict__ z, const double *done)
{
    if (done && *done != 0.0) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) z[i] = (double)x[i];
}

// x = A_c^-1 b (dense, one 64-lane wave per row)
__global__ __launch_bounds__(64) void k_mg_coarse(int64_t n3, const double *__restrict__ ainv,
                                                  const double *__restrict__ b, double *__restrict__ x,
                                                  const double *done)
{
    if (done && *done != 0.0) return;
    const int64_t row = blockIdx.x;
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n3; k += 64) s += ainv[row * n3 + k] * b[k];
    s = wave_sum64(s);
    if (threadIdx.x == 0) x[row] = s;
}

// ---- host helpers ------------------------------------------------------------------------------------
template <typename T>
int upload(ewfe_mg *mg, T *&dst, const std::vector<T> &src)
{
    dst = nullptr;
    if (src.empty()) return 0;
    HIP_CHECK(hipMalloc((void **)&dst, sizeof(T) * src.size()));
    mg->allocs.push_back((void *)dst);
    HIP_CHECK(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int alloc(ewfe_mg *mg, T *&dst, size_t count)
{
    dst = nullptr;
    if (count == 0) return 0;
    HIP_CHECK(hipMalloc((void **)&dst, sizeof(T) * count));
    mg->allocs.push_back((void *)dst);
    return 0;
}

#define RC(expr)                 \
    do {                         \
        const int _rc = (expr);  \
        if (_rc) return _rc;     \
    } while (0)

template <typename VT>
FineOp<VT> fine_op(const ewfe_mg *mg, const void *vals)
{
    const ewfe_plan *p = mg->plan;
    FineOp<VT> F;
    F.meta = reinterpret_cast<const int4 *>(p->d_rn_meta);
    F.nbcol = p->d_nbcol;
    F.diag_rank = p->d_diag_rank;
    F.vals = (const VT *)vals;
    F.mask = mg->mask;
    F.n = p->n_rownodes;
    return F;
}

BsrOp bsr_op(const MgLevel &L)
{
    BsrOp B;
    B.aptr = L.aptr;
    B.acol = L.acol;
    B.adiag = L.adiag;
    B.aval = L.aval;
    B.n = L.n;
    return B;
}

// lambda_max(D^-1 A_l) into mg->lam[l]
template <typename OP>
int power_iteration(ewfe_mg *mg, int l, const OP &A, const double *dinv, const uint8_t *mask, hipStream_t s)
{
    const int64_t n3 = 3 * A.n;
    hipLaunchKernelGGL(k_mg_power_init, dim3(1024), dim3(256), 0, s, n3, mask, mg->pw_x);
    for (int it = 0; it < kPowerIters; ++it) {
        hipLaunchKernelGGL((k_mg_power<OP>), dim3(kBlocks), dim3(kThreads), 0, s, A, dinv, (const double *)mg->pw_x,
                           mg->pw_y, mg->partials);
        hipLaunchKernelGGL(k_mg_norm, dim3(1), dim3(1024), 0, s, (const double *)mg->partials, mg->lam + l);
        if (it + 1 < kPowerIters)
            hipLaunchKernelGGL(k_mg_scale, dim3(1024), dim3(256), 0, s, n3, (const double *)mg->pw_y, mg->pw_x,
                               (const double *)(mg->lam + l));
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

// dense inverse of the coarsest operator (host, Cholesky; zero rows/columns get a unit diagonal)
int coarse_inverse(ewfe_mg *mg, hipStream_t s)
{
    const MgLevel &C = mg->lv.back();
    const int64_t n3 = 3 * C.n;
    std::vector<double> av(9 * (size_t)C.nnzb);
    HIP_CHECK(hipMemcpyAsync(av.data(), C.aval, sizeof(double) * av.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> A((size_t)n3 * n3, 0.0);
    for (int64_t I = 0; I < C.n; ++I)
        for (int32_t e = mg->c_aptr[I]; e < mg->c_aptr[I + 1]; ++e) {
            const int64_t J = mg->c_acol[e];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) A[(3 * I + r) * n3 + 3 * J + c] = av[9 * (size_t)e + 3 * r + c];
        }
    for (int64_t k = 0; k < n3; ++k) {
        bool zero = true;
        for (int64_t t = 0; t < n3 && zero; ++t) zero = A[k * n3 + t] == 0.0 && A[t * n3 + k] == 0.0;
        if (zero) A[k * n3 + k] = 1.0;
    }
    // symmetrise (A_c is symmetric up to rounding), Cholesky A = L L^T in the lower triangle
    for (int64_t i = 0; i < n3; ++i)
        for (int64_t j = 0; j < i; ++j) {
            const double v = 0.5 * (A[i * n3 + j] + A[j * n3 + i]);
            A[i * n3 + j] = A[j * n3 + i] = v;
        }
    for (int64_t j = 0; j < n3; ++j) {
        double dg = A[j * n3 + j];
        for (int64_t k = 0; k < j; ++k) dg -= A[j * n3 + k] * A[j * n3 + k];
        if (!(dg > 0.0)) {
            ewfe::set_error("multigrid: coarsest operator is not positive definite");
            return EWFE_EINVAL;
        }
        const double ljj = std::sqrt(dg);
        A[j * n3 + j] = ljj;
        for (int64_t i = j + 1; i < n3; ++i) {
            double v = A[i * n3 + j];
            for (int64_t k = 0; k < j; ++k) v -= A[i * n3 + k] * A[j * n3 + k];
            A[i * n3 + j] = v / ljj;
        }
    }
    // M = L^-1 (lower triangular), row by row with contiguous row updates: M_i = -(sum_k<i L_ik M_k) / L_ii + e_i/L_ii
    std::vector<double> Mi((size_t)n3 * n3, 0.0), tmp(n3);
    for (int64_t i = 0; i < n3; ++i) {
        std::fill(tmp.begin(), tmp.begin() + i + 1, 0.0);
        for (int64_t k = 0; k < i; ++k) {
            const double l = A[i * n3 + k];
            if (l == 0.0) continue;
            const double *mk = Mi.data() + k * n3;
            for (int64_t j = 0; j <= k; ++j) tmp[j] += l * mk[j];
        }
        const double dinv = 1.0 / A[i * n3 + i];
        double *mi = Mi.data() + i * n3;
        for (int64_t j = 0; j < i; ++j) mi[j] = -tmp[j] * dinv;
        mi[i] = dinv;
    }
    // A^-1 = M^T M = sum_k (row k of M)^T (row k of M), lower triangle accumulated, then mirrored
    std::vector<double> inv((size_t)n3 * n3, 0.0);
    for (int64_t k = 0; k < n3; ++k) {
        const double *mk = Mi.data() + k * n3;
        for (int64_t i = 0; i <= k; ++i) {
            const double a = mk[i];
            if (a == 0.0) continue;
            double *ri = inv.data() + i * n3;
            for (int64_t j = 0; j <= i; ++j) ri[j] += a * mk[j];
        }
    }
    for (int64_t i = 0; i < n3; ++i)
        for (int64_t j = 0; j < i; ++j) inv[j * n3 + i] = inv[i * n3 + j];
    HIP_CHECK(hipMemcpyAsync(mg->ainv, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

template <typename VT>
int mg_setup(ewfe_mg *mg, const void *vals, const int32_t *rows, int64_t n_rows, hipStream_t s)
{
    const ewfe_plan *p = mg->plan;
    HIP_CHECK(hipMemsetAsync(mg->mask, 0, (size_t)p->n_dof, s));
    if (n_rows > 0) hipLaunchKernelGGL(k_mg_mask, dim3(256), dim3(256), 0, s, rows, n_rows, mg->mask);
    const FineOp<VT> F = fine_op<VT>(mg, vals);
    const int nl = (int)mg->lv.size();
    for (int l = 0; l < nl; ++l) {
        MgLevel &L = mg->lv[l];
        const BsrOp B = bsr_op(L);
        if (l == 0) {
            hipLaunchKernelGGL((k_mg_dinv_fine<VT>), dim3(1024), dim3(256), 0, s, F, L.dinv);
            if constexpr (std::is_same<VT, float>::value) {
                if (!mg->dinv32) RC(alloc(mg, mg->dinv32, 9 * (size_t)L.n));
                hipLaunchKernelGGL(k_mg_narrow, dim3(1024), dim3(256), 0, s, 9 * L.n, (const double *)L.dinv,
                                   mg->dinv32);
            }
            RC(power_iteration(mg, 0, F, L.dinv, mg->mask, s));
        } else {
            hipLaunchKernelGGL(k_mg_dinv_bsr, dim3(1024), dim3(256), 0, s, B, L.dinv);
            RC(power_iteration(mg, l, B, L.dinv, (const uint8_t *)nullptr, s));
        }
        if (l + 1 == nl) break;
        MgLevel &N = mg->lv[l + 1];
        if (l == 0) {
            hipLaunchKernelGGL((k_mg_prolongator<true, VT, VT>), dim3((unsigned)L.n), dim3(64), 0, s, F, B,
                               (const double *)L.dinv, (const int32_t *)L.pptr, (const uint8_t *)L.pslot,
                               (VT *)L.pval, (const double *)(mg->lam + l));
            const int64_t chunk = (L.n + 7) / 8;
            if (mg->plan->max_deg <= kFineMaxRow)
                hipLaunchKernelGGL((k_mg_ap<true, VT, VT, kFineMaxRow, kFineMaxStage, false>),
                                   dim3((unsigned)(8 * chunk)), dim3(64), 0, s, F, B, (const int32_t *)L.pptr,
                                   (const int32_t *)L.pcol, (const VT *)L.pval, (const int32_t *)L.wptr,
                                   (const int32_t *)L.wcol, L.wval, L.n, chunk);
            else
                hipLaunchKernelGGL((k_mg_ap<true, VT, VT, kWideMaxRow, kWideMaxStage, false>),
                                   dim3((unsigned)(8 * chunk)), dim3(64), 0, s, F, B, (const int32_t *)L.pptr,
                                   (const int32_t *)L.pcol, (const VT *)L.pval, (const int32_t *)L.wptr,
                                   (const int32_t *)L.wcol, L.wval, L.n, chunk);
        } else {
            hipLaunchKernelGGL((k_mg_prolongator<false, double, VT>), dim3((unsigned)L.n), dim3(64), 0, s,
                               FineOp<double>{}, B, (const double *)L.dinv, (const int32_t *)L.pptr,
                               (const uint8_t *)L.pslot, (VT *)L.pval, (const double *)(mg->lam + l));
            const int64_t chunk = (L.n + 7) / 8;
            hipLaunchKernelGGL((k_mg_ap<false, double, VT, kCoarseMaxRow, kCoarseMaxStage, true>),
                               dim3((unsigned)(8 * chunk)), dim3(64), 0, s, FineOp<double>{}, B,
                               (const int32_t *)L.pptr, (const int32_t *)L.pcol, (const VT *)L.pval,
                               (const int32_t *)L.wptr, (const int32_t *)L.wcol, L.wval, L.n, chunk);
        }
        hipLaunchKernelGGL(k_mg_rap<VT>, dim3((unsigned)N.n), dim3(64), 0, s, N.n, (const int32_t *)L.rptr,
                           (const int32_t *)L.ri, (const int32_t *)L.rt, (const VT *)L.pval,
                           (const int32_t *)L.wptr, (const int32_t *)L.wcol, (const double *)L.wval,
                           (const int32_t *)N.aptr, (const int32_t *)N.acol, N.aval);
        if (L.n_r > 0)
            hipLaunchKernelGGL(k_mg_rgather<VT>, dim3(2048), dim3(256), 0, s, L.n_r, (const int32_t *)L.rt,
                               (const VT *)L.pval, (VT *)L.rval);
        HIP_CHECK(hipGetLastError());
    }
    RC(coarse_inverse(mg, s));
    if (mg->fine16) {
        const int64_t nnz = p->nnz;
        hipLaunchKernelGGL(k_mg_zero2, dim3(1), dim3(64), 0, s, mg->scale16);
        hipLaunchKernelGGL(k_mg_absmax<VT>, dim3(2048), dim3(256), 0, s, (const VT *)vals, nnz, mg->scale16);
        hipLaunchKernelGGL(k_mg_to_half<VT>, dim3(4096), dim3(256), 0, s, (const VT *)vals, nnz, mg->scale16,
                           mg->vals16);
        HIP_CHECK(hipGetLastError());
    }
    mg->ready = true;
    mg->setup_dtype = std::is_same<VT, float>::value ? EWFE_F32 : EWFE_F64;
    return 0;
}

// one V-cycle on level l: x_l = M_l^-1 b_l
template <typename VT>
int vcycle(ewfe_mg *mg, int l, const void *vals, const double *done, hipStream_t s)
{
    MgLevel &L = mg->lv[l];
    const int nl = (int)mg->lv.size();
    if (l + 1 == nl) {
        hipLaunchKernelGGL(k_mg_coarse, dim3((unsigned)(3 * L.n)), dim3(64), 0, s, 3 * L.n, (const double *)mg->ainv,
                           (const double *)L.b, L.x, done);
        return 0;
    }
    MgLevel &N = mg->lv[l + 1];
    const double *lam = mg->lam + l;
    const double ratio = mg->ratio;
    const int nu = mg->degree;
    const unsigned ew = 1024;  // grid of the node-wise kernels
    // pre-smoothing from x = 0, residual, restriction, the next level, prolongation, post-smoothing; vectors r, x, d of
    // type XT (b and the next level's vectors are FP64)
    auto run_level = [&](auto &&op, const auto *dinv, auto *r, auto *x, auto *da, auto *db) -> int {
        using OP = std::decay_t<decltype(op)>;
        using DT = std::remove_const_t<std::remove_pointer_t<decltype(dinv)>>;
        using XT = std::remove_pointer_t<decltype(r)>;
        XT *dc = da, *dn = db;
        auto smooth_steps = [&]() {
            for (int k = 1; k < nu; ++k) {
                hipLaunchKernelGGL((k_mg_cheb_step<OP, DT, XT>), dim3(kBlocks), dim3(kThreads), 0, s, op, dinv, r,
                                   (const XT *)dc, dn, x, lam, ratio, k, done);
                std::swap(dc, dn);
            }
        };
        hipLaunchKernelGGL((k_mg_cheb0_zero<DT, XT>), dim3(ew), dim3(256), 0, s, L.n, dinv, (const double *)L.b, r,
                           dc, x, lam, ratio, done);
        smooth_steps();
        // residual of the smoothed iterate: r -= A d_last
        hipLaunchKernelGGL((k_mg_resid<OP, XT, XT>), dim3(kBlocks), dim3(kThreads), 0, s, op, (const XT *)r,
                           (const XT *)dc, r, done);
        hipLaunchKernelGGL((k_mg_restrict<VT, XT>), dim3((unsigned)N.n), dim3(64), 0, s, N.n, (const int32_t *)L.rptr,
                           (const int32_t *)L.ri, (const VT *)L.rval, (const XT *)r, N.b,
                           done);
        RC(vcycle<VT>(mg, l + 1, vals, done, s));
        hipLaunchKernelGGL((k_mg_prolong<VT, XT>), dim3(ew), dim3(256), 0, s, L.n, (const int32_t *)L.pptr,
                           (const int32_t *)L.pcol, (const VT *)L.pval, (const double *)N.x, x, done);
        dc = da, dn = db;
        hipLaunchKernelGGL((k_mg_resid<OP, double, XT>), dim3(kBlocks), dim3(kThreads), 0, s, op,
                           (const double *)L.b, (const XT *)x, r, done);
        hipLaunchKernelGGL((k_mg_cheb0<DT, XT>), dim3(ew), dim3(256), 0, s, L.n, dinv, (const XT *)r, dc, x, lam,
                           ratio, done);
        smooth_steps();
        return 0;
    };
    if constexpr (std::is_same<VT, float>::value) {
        if (l == 0) {
            // FP32 matrix: the fine level's vectors and block-Jacobi inverses in FP32 too (the vectors inside the
            // level's own buffers, 3n doubles each: r and x share the first), the caller's b read and its result
            // written in FP64
            float *rf = reinterpret_cast<float *>(L.r), *xf = rf + 3 * L.n;
            float *da = reinterpret_cast<float *>(L.d0), *db = reinterpret_cast<float *>(L.d1);
            const float *dinv = mg->dinv32;
            if (mg->fine16) {  // the fine-level passes on the half-precision copy of the set-up matrix
                FineOp<_Float16> F = fine_op<_Float16>(mg, mg->vals16);
                F.scale = mg->scale16;
                RC(run_level(F, dinv, rf, xf, da, db));
            } else {
                RC(run_level(fine_op<VT>(mg, vals), dinv, rf, xf, da, db));
            }
            hipLaunchKernelGGL(k_mg_widen, dim3(ew), dim3(256), 0, s, 3 * L.n, (const float *)xf, L.x, done);
            HIP_CHECK(hipGetLastError());
            return 0;
        }
    }
    if (l == 0 && mg->fine16) {
        FineOp<_Float16> F = fine_op<_Float16>(mg, mg->vals16);
        F.scale = mg->scale16;
        RC(run_level(F, (const double *)L.dinv, L.r, L.x, L.d0, L.d1));
    } else if (l == 0) {
        RC(run_level(fine_op<VT>(mg, vals), (const double *)L.dinv, L.r, L.x, L.d0, L.d1));
    } else {
        RC(run_level(bsr_op(L), (const double *)L.dinv, L.r, L.x, L.d0, L.d1));
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

namespace ewfe {

int mg_apply(ewfe_mg *mg, const void *vals, int vals_dtype, const double *r, double *z, const double *done,
             void *stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (vals_dtype != mg->setup_dtype) {
        ewfe::set_error("multigrid: the hierarchy was set up for a matrix of another precision");
        return EWFE_EINVAL;
    }
    MgLevel &L0 = mg->lv[0];
    L0.b = const_cast<double *>(r);
    L0.x = z;
    if (vals_dtype == EWFE_F64) return vcycle<double>(mg, 0, vals, done, s);
    return vcycle<float>(mg, 0, vals, done, s);
}

bool mg_ready(const ewfe_mg *mg) { return mg && mg->ready; }
const ewfe_plan *mg_plan(const ewfe_mg *mg) { return mg ? mg->plan : nullptr; }

}  // namespace ewfe

extern "C" void ewfe_mg_destroy(ewfe_mg *mg)
{
    if (!mg) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(mg->device);
    for (void *q : mg->allocs) (void)hipFree(q);
    (void)hipSetDevice(prev);
    delete mg;
}

extern "C" int ewfe_mg_create(const ewfe_plan *p, int degree, double ratio, ewfe_mg **out, int *n_levels_out)
{
    if (!p || !out || degree < 1 || degree > 8 || !(ratio > 1.0)) {
        ewfe::set_error("ewfe_mg_create: invalid arguments");
        return EWFE_EINVAL;
    }
    *out = nullptr;
    if (!p->rows_cover_all_dofs || !p->d_rn_meta || !p->d_nbcol || p->n_rownodes * 3 != p->n_dof) {
        ewfe::set_error("ewfe_mg_create: the multigrid needs 3 dofs on every node of the plan");
        return EWFE_EINVAL;
    }
    if (p->max_deg > kAmgMaxRow) {
        ewfe::set_error("ewfe_mg_create: a node has more than " + std::to_string(ewfe::kAmgMaxRow) + " neighbours");
        return EWFE_EINVAL;
    }
    // fine node graph from the plan's neighbour lists (row node r owns dofs 3r .. 3r+2)
    const int64_t n = p->n_rownodes;
    std::vector<int32_t> meta(4 * (size_t)n), nbcol(p->n_nbblocks);
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(p->device));
    HIP_CHECK(hipMemcpy(meta.data(), p->d_rn_meta, sizeof(int32_t) * meta.size(), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(nbcol.data(), p->d_nbcol, sizeof(int32_t) * nbcol.size(), hipMemcpyDeviceToHost));
    std::vector<int32_t> aptr(n + 1), acol(p->n_nbblocks);
    for (int64_t i = 0; i < n; ++i) {
        if (meta[4 * i] != 3 * i) {
            (void)hipSetDevice(prev);
            ewfe::set_error("ewfe_mg_create: row nodes do not own consecutive dof triples");
            return EWFE_EINVAL;
        }
        aptr[i] = meta[4 * i + 2];
        aptr[i + 1] = meta[4 * i + 2] + meta[4 * i + 3];
        for (int32_t e = aptr[i]; e < aptr[i + 1]; ++e) acol[e] = nbcol[e] / 3;
    }
    ewfe::AmgHost H;
    int rc = ewfe::amg_build_host(n, aptr, acol, 200, 10, H);
    if (rc) {
        (void)hipSetDevice(prev);
        return rc;
    }
    ewfe_mg *mg = new ewfe_mg;
    mg->plan = p;
    mg->device = p->device;
    mg->degree = degree;
    mg->ratio = ratio;
    auto fail = [&](int code) {
        ewfe_mg_destroy(mg);
        (void)hipSetDevice(prev);
        return code;
    };
    const int nl = (int)H.levels.size();
    mg->lv.resize(nl);
    for (int l = 0; l < nl; ++l) {
        const ewfe::AmgLevelHost &h = H.levels[l];
        MgLevel &L = mg->lv[l];
        L.n = h.n;
        L.nnzb = h.aptr[h.n];
        if (l > 0) {
            if ((rc = upload(mg, L.aptr, h.aptr)) || (rc = upload(mg, L.acol, h.acol)) ||
                (rc = upload(mg, L.adiag, h.adiag)) || (rc = alloc(mg, L.aval, 9 * (size_t)L.nnzb)))
                return fail(rc);
        }
        if ((rc = alloc(mg, L.dinv, 9 * (size_t)L.n))) return fail(rc);
        if (l > 0) {
            if ((rc = alloc(mg, L.b, 3 * (size_t)L.n)) || (rc = alloc(mg, L.x, 3 * (size_t)L.n))) return fail(rc);
        }
        if ((rc = alloc(mg, L.r, 3 * (size_t)L.n)) || (rc = alloc(mg, L.d0, 3 * (size_t)L.n)) ||
            (rc = alloc(mg, L.d1, 3 * (size_t)L.n)))
            return fail(rc);
        if (l + 1 < nl) {
            L.n_agg = h.n_agg;
            if ((rc = upload(mg, L.pptr, h.pptr)) || (rc = upload(mg, L.pcol, h.pcol)) ||
                (rc = upload(mg, L.pslot, h.pslot)) || (rc = alloc(mg, L.pval, 9 * h.pcol.size())) ||
                (rc = upload(mg, L.wptr, h.wptr)) || (rc = upload(mg, L.wcol, h.wcol)) ||
                (rc = alloc(mg, L.wval, 9 * h.wcol.size())) || (rc = upload(mg, L.rptr, h.rptr)) ||
                (rc = upload(mg, L.ri, h.ri)) || (rc = upload(mg, L.rt, h.rt)) ||
                (rc = alloc(mg, L.rval, 9 * h.rt.size())))
                return fail(rc);
            L.n_r = (int64_t)h.rt.size();
            // k_mg_ap keeps a row of A_l in LDS
            const int64_t max_row = l > 0 ? kCoarseMaxRow : (p->max_deg <= kFineMaxRow ? kFineMaxRow : kWideMaxRow);
            for (int64_t i = 0; i < h.n; ++i)
                if (h.aptr[i + 1] - h.aptr[i] > max_row) {
                    ewfe::set_error("ewfe_mg_create: level " + std::to_string(l) + " has a row with more than " +
                                    std::to_string(max_row) + " blocks");
                    return fail(EWFE_EINVAL);
                }
        }
    }
    mg->c_aptr = H.levels.back().aptr;
    mg->c_acol = H.levels.back().acol;
    mg->nc3 = 3 * H.levels.back().n;
    if ((rc = alloc(mg, mg->ainv, (size_t)mg->nc3 * mg->nc3)) || (rc = alloc(mg, mg->lam, (size_t)nl)) ||
        (rc = alloc(mg, mg->partials, (size_t)kBlocks)) || (rc = alloc(mg, mg->mask, (size_t)p->n_dof)) ||
        (rc = alloc(mg, mg->pw_x, 3 * (size_t)n)) || (rc = alloc(mg, mg->pw_y, 3 * (size_t)n)))
        return fail(rc);
    (void)hipSetDevice(prev);
    *out = mg;
    if (n_levels_out) *n_levels_out = nl;
    return 0;
}

extern "C" int ewfe_mg_level_info(const ewfe_mg *mg, int level, int64_t *n_nodes, int64_t *nnz_blocks)
{
    if (!mg || level < 0 || level >= (int)mg->lv.size()) {
        ewfe::set_error("ewfe_mg_level_info: invalid arguments");
        return EWFE_EINVAL;
    }
    if (n_nodes) *n_nodes = mg->lv[level].n;
    if (nnz_blocks) *nnz_blocks = level == 0 ? mg->plan->n_nbblocks : mg->lv[level].nnzb;
    return 0;
}

extern "C" int ewfe_mg_setup(ewfe_mg *mg, const void *vals, int vals_dtype, const int32_t *rows_dev, int64_t n_rows,
                             void *stream)
{
    if (!mg || !vals || (vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32 && vals_dtype != EWFE_F32_D64) ||
        (n_rows > 0 && !rows_dev)) {
        ewfe::set_error("ewfe_mg_setup: invalid arguments");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    mg->ready = false;
    if (vals_dtype == EWFE_F64) return mg_setup<double>(mg, vals, rows_dev, n_rows, s);
    // EWFE_F32_D64: the hierarchy is built from the FP32 values in front of the FP64 diagonal blocks
    return mg_setup<float>(mg, vals, rows_dev, n_rows, s);
}

extern "C" int ewfe_mg_set_fine_precision(ewfe_mg *mg, int dtype)
{
    if (!mg || (dtype != EWFE_F16 && dtype != EWFE_F32 && dtype != EWFE_F64)) {
        ewfe::set_error("ewfe_mg_set_fine_precision: invalid arguments");
        return EWFE_EINVAL;
    }
    const bool want = dtype == EWFE_F16;
    if (want && !mg->vals16) {
        int prev = 0;
        HIP_CHECK(hipGetDevice(&prev));
        HIP_CHECK(hipSetDevice(mg->device));
        int rc = alloc(mg, mg->vals16, (size_t)mg->plan->nnz);
        if (!rc) rc = alloc(mg, mg->scale16, (size_t)2);
        (void)hipSetDevice(prev);
        if (rc) return rc;
    }
    if (want != mg->fine16) mg->ready = false;  // the copy is made by the next ewfe_mg_setup
    mg->fine16 = want;
    return 0;
}

extern "C" int ewfe_mg_lambda(const ewfe_mg *mg, double *lam_out, int n)
{
    if (!mg || !lam_out || n < (int)mg->lv.size()) {
        ewfe::set_error("ewfe_mg_lambda: invalid arguments");
        return EWFE_EINVAL;
    }
    HIP_CHECK(hipMemcpy(lam_out, mg->lam, sizeof(double) * mg->lv.size(), hipMemcpyDeviceToHost));
    return 0;
}
