// libewfe.so -- element evaluation + CSR assembly kernels for gfx950 (MI355X), and the C ABI.
//
// What one ewfe_eval_assemble call computes is the reference's per-iteration element loop
// (edelweissfe/solvers/nonlinearimplicitstaticparallelmk2.pyx:111-198 -> MarmotElement::computeYourself)
// followed by CSRGenerator.updateCSR (edelweissfe/numerics/csrgenerator.pyx:100-115):
//   P = sum_e Pe (Pe = -f_int,e), F = sum_e |Pe|, CSR values = sum_e Ke scattered through the slot map.
// The element formulations follow the in-tree specs of the reference:
//   small strain: edelweissfe/elements/displacementelement/element.py:290-346
//   total Lagrangian: edelweissfe/elements/displacementtlelement/element.py:346-427
//   J2: edelweissfe/materials/vonmises/vonmises.py:186-254; linear elastic: linearelastic.py:185-210
//   Neo-Hooke Wa: edelweissfe/materials/neohooke/neohookepencegouformulationa.py:107-141
//
// MI355X design
//  * C3D8 family, two passes. Pass 1 (k_hex8_elem): a 64-lane workgroup holds 8 elements. Nodal X/U/dU are
//    staged in LDS (one node per lane), one quadrature point per lane runs the Jacobian, gradients and the
//    material update, forms its contributions to the nodal forces (summed over the element's 8 lanes in a fixed
//    order) and leaves the per-point tangent "recipe" in LDS, in the precision of the records; then one node per
//    lane builds the 3x3 blocks K_ab for b in {a, a+1, a+2, a+3 (, a+4 if a<4)} mod 8 -- the 36 blocks of the
//    symmetric 24x24 Ke, each exactly once -- and the element's record (36 blocks + nodal forces) is written
//    contiguously. Pass 2 (k_gather8s on boxGen grids, k_gather8 elsewhere): per row node, the blocks of its
//    incident elements are summed in ascending element order and the node's 3 CSR rows, P and F are written once.
//    Deterministic FP64 without atomics: every output is one lane's sum in a fixed order.
//  * Quadrature-point state is structure-of-arrays [component][qp][plan position], so consecutive lanes
//    touch consecutive addresses.
//  * The tangent uses the rank structure of the isotropic/J2/Neo-Hooke tangents instead of a dense
//    B^T C B:  K_ab,ik = sum_q [ alpha p_a,i p_b,k + beta p_a,k p_b,i + gamma delta_ik (r_a . r_b)
//                                - c v_a,i v_b,k ] ,  v = N p
//    with p = r = dN/dx (small strain) or p = dN/dx (spatial), r = dN/dX (total Lagrangian).
//  * C3D20 family (k_hex20): one element per 64-lane workgroup, four lanes per quadrature point, then lanes
//    (node a, group) build the blocks of their node pairs; the element's record is written for the row-node
//    gather (k_gather20, the same ascending-element sums as pass 2 above). Meshes beyond the gather's limits take
//    the colour path instead: colour by colour (no two elements of a colour share a node), each element adds its
//    blocks into the CSR values, which fixes the order of every sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "ewfe.h"
#include "ewfe_internal.h"

// ------------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------------
static thread_local std::string g_err;
void ewfe::set_error(const std::string &m) { g_err = m; }
extern "C" const char *ewfe_last_error(void) { return g_err.c_str(); }
extern "C" int ewfe_version(void) { return 1; }

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            ewfe::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                  \
            return EWFE_EHIP;                                                                    \
        }                                                                                        \
    } while (0)

// ------------------------------------------------------------------------------------------------
// shape-function tables (Abaqus node order; Gauss points q = i + n*j + n*n*k over (xi, eta, zeta))
// ------------------------------------------------------------------------------------------------
__constant__ double c_dN8[8][8][3];
__constant__ double c_w8[8];
__constant__ double c_dN20[27][20][3];
__constant__ double c_w27[27];
__constant__ double c_N8[8][8];    // shape function values at the quadrature points
__constant__ double c_N20[27][20];
__constant__ double c_dN20r[8][20][3];  // 20-node functions at the 2x2x2 points (C3D20R)
__constant__ double c_N20r[8][20];

// the element's tables by node count NN and point count NQ
typedef const double (*DNRow)[3];
template <int NN, int NQ>
__device__ __forceinline__ DNRow qp_dN(int q)
{
    return NN == 8 ? c_dN8[q] : (NQ == 8 ? c_dN20r[q] : c_dN20[q]);
}
template <int NN, int NQ>
__device__ __forceinline__ const double *qp_N(int q)
{
    return NN == 8 ? c_N8[q] : (NQ == 8 ? c_N20r[q] : c_N20[q]);
}
template <int NN, int NQ>
__device__ __forceinline__ double qp_w(int q)
{
    return NQ == 8 ? c_w8[q] : c_w27[q];
}

static const double kHexXi[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {0, -1, 1},  {1, 0, 1},  {0, 1, 1}, {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0}};

// signs of eta and zeta of the 2x2x2 points, by q mod 4 (the kernels use the same signs as bit arithmetic:
// eta = bit 1 of q, zeta = bit 0 xor bit 1)
static const double kEta8[4] = {-1.0, -1.0, 1.0, 1.0};
static const double kZeta8[4] = {-1.0, 1.0, 1.0, -1.0};

static void host_dN(int nn, const double x[3], double dN[][3])
{
    for (int a = 0; a < nn; ++a) {
        const double xa = kHexXi[a][0], ya = kHexXi[a][1], za = kHexXi[a][2];
        if (nn == 8) {
            dN[a][0] = 0.125 * xa * (1 + ya * x[1]) * (1 + za * x[2]);
            dN[a][1] = 0.125 * (1 + xa * x[0]) * ya * (1 + za * x[2]);
            dN[a][2] = 0.125 * (1 + xa * x[0]) * (1 + ya * x[1]) * za;
        } else if (a < 8) {  // serendipity corner: 1/8 (1+xa x)(1+ya y)(1+za z)(xa x + ya y + za z - 2)
            const double fx = 1 + xa * x[0], fy = 1 + ya * x[1], fz = 1 + za * x[2];
            const double s = xa * x[0] + ya * x[1] + za * x[2] - 2;
            dN[a][0] = 0.125 * xa * fy * fz * (s + fx);
            dN[a][1] = 0.125 * ya * fx * fz * (s + fy);
            dN[a][2] = 0.125 * za * fx * fy * (s + fz);
        } else {  // mid-side: 1/4 (1 - t^2) along the zero axis
            double f[3], df[3];
            const double c[3] = {xa, ya, za};
            for (int d = 0; d < 3; ++d) {
                if (c[d] == 0) {
                    f[d] = 1 - x[d] * x[d];
                    df[d] = -2 * x[d];
                } else {
                    f[d] = 1 + c[d] * x[d];
                    df[d] = c[d];
                }
            }
            dN[a][0] = 0.25 * df[0] * f[1] * f[2];
            dN[a][1] = 0.25 * f[0] * df[1] * f[2];
            dN[a][2] = 0.25 * f[0] * f[1] * df[2];
        }
    }
}

// Shape functions and parametric derivatives at an arbitrary point (device), same node tables as below
__constant__ double c_hexXi[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {0, -1, 1},  {1, 0, 1},  {0, 1, 1}, {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0}};
template <int NN>
__device__ void shape8_20(const double x[3], double N[NN], double dN[NN][3])
{
    for (int a = 0; a < NN; ++a) {
        const double c[3] = {c_hexXi[a][0], c_hexXi[a][1], c_hexXi[a][2]};
        if (NN == 8) {
            const double f0 = 1 + c[0] * x[0], f1 = 1 + c[1] * x[1], f2 = 1 + c[2] * x[2];
            N[a] = 0.125 * f0 * f1 * f2;
            dN[a][0] = 0.125 * c[0] * f1 * f2;
            dN[a][1] = 0.125 * f0 * c[1] * f2;
            dN[a][2] = 0.125 * f0 * f1 * c[2];
        } else if (a < 8) {
            const double f0 = 1 + c[0] * x[0], f1 = 1 + c[1] * x[1], f2 = 1 + c[2] * x[2];
            const double s = c[0] * x[0] + c[1] * x[1] + c[2] * x[2] - 2;
            N[a] = 0.125 * f0 * f1 * f2 * s;
            dN[a][0] = 0.125 * c[0] * f1 * f2 * (s + f0);
            dN[a][1] = 0.125 * c[1] * f0 * f2 * (s + f1);
            dN[a][2] = 0.125 * c[2] * f0 * f1 * (s + f2);
        } else {
            double f[3], df[3];
            for (int d = 0; d < 3; ++d) {
                if (c[d] == 0) {
                    f[d] = 1 - x[d] * x[d];
                    df[d] = -2 * x[d];
                } else {
                    f[d] = 1 + c[d] * x[d];
                    df[d] = c[d];
                }
            }
            N[a] = 0.25 * f[0] * f[1] * f[2];
            dN[a][0] = 0.25 * df[0] * f[1] * f[2];
            dN[a][1] = 0.25 * f[0] * df[1] * f[2];
            dN[a][2] = 0.25 * f[0] * f[1] * df[2];
        }
    }
}

// Shape function values: trilinear (8 nodes) and serendipity (20 nodes, Abaqus node order)
static void host_N(int nn, const double x[3], double N[])
{
    for (int a = 0; a < nn; ++a) {
        const double xa = kHexXi[a][0], ya = kHexXi[a][1], za = kHexXi[a][2];
        if (nn == 8) {
            N[a] = 0.125 * (1 + xa * x[0]) * (1 + ya * x[1]) * (1 + za * x[2]);
        } else if (a < 8) {
            N[a] = 0.125 * (1 + xa * x[0]) * (1 + ya * x[1]) * (1 + za * x[2]) * (xa * x[0] + ya * x[1] + za * x[2] - 2);
        } else {
            const double c[3] = {xa, ya, za};
            double f = 0.25;
            for (int d = 0; d < 3; ++d) f *= (c[d] == 0) ? (1 - x[d] * x[d]) : (1 + c[d] * x[d]);
            N[a] = f;
        }
    }
}

static int upload_tables(int device)
{
    static std::mutex mu;
    static bool done[256] = {false};
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0 || device >= 256) return EWFE_EINVAL;
    if (done[device]) return 0;
    double dN8[8][8][3], w8[8], dN20[27][20][3], w27[27], N8[8][8], N20[27][20], dN20r[8][20][3], N20r[8][20];
    const double g2 = 1.0 / std::sqrt(3.0), g3 = std::sqrt(0.6);
    // quadrature points in the reference's order (library.py:212-227, 260-275): 2x2x2: xi = -g for q < 4,
    // +g after, (eta, zeta) = (-,-), (-,+), (+,+), (+,-) by q mod 4; 3x3x3: q = 9 i + 3 j + k over (xi, eta, zeta)
    for (int q = 0; q < 8; ++q) {
        const double x[3] = {q < 4 ? -g2 : g2, kEta8[q & 3] * g2, kZeta8[q & 3] * g2};
        host_dN(8, x, dN8[q]);
        host_N(8, x, N8[q]);
        host_dN(20, x, dN20r[q]);
        host_N(20, x, N20r[q]);
        w8[q] = 1.0;
    }
    const double pts[3] = {-g3, 0.0, g3}, wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int q = 0; q < 27; ++q) {
        const int i = q / 9, j = (q / 3) % 3, k = q % 3;
        const double x[3] = {pts[i], pts[j], pts[k]};
        host_dN(20, x, dN20[q]);
        host_N(20, x, N20[q]);
        w27[q] = wts[i] * wts[j] * wts[k];
    }
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_N8), N8, sizeof(N8)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_N20), N20, sizeof(N20)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dN8), dN8, sizeof(dN8)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_w8), w8, sizeof(w8)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dN20), dN20, sizeof(dN20)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_w27), w27, sizeof(w27)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_dN20r), dN20r, sizeof(dN20r)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_N20r), N20r, sizeof(N20r)));
    done[device] = true;
    return 0;
}

// ------------------------------------------------------------------------------------------------
// per-quadrature-point material update -> stress S (wd folded in later) + tangent recipe
// ------------------------------------------------------------------------------------------------
struct MatParams {
    double p[8];
};

enum { FORM_SS = 0, FORM_TL = 1 };
enum { M_LE = 0, M_J2 = 1, M_NH = 2, M_SVK = 3 };

struct Recipe {
    double S[6];  // stress, tensor components [11,22,33,12,13,23]
    double alpha, beta, gam, c;
    double N[6];  // J2 flow direction (tensor, unit norm), used when c != 0
};

__device__ __forceinline__ double det3(const double A[3][3])
{
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

__device__ __forceinline__ double inv3(const double A[3][3], double B[3][3])
{
    const double d = det3(A);
    const double r = 1.0 / d;
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * r;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * r;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * r;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * r;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
    return d;
}

// Linear elasticity (linearelastic.py:95-117): lambda = f nu, mu = f (1 - 2 nu) / 2,
// f = E / ((1 + nu)(1 - 2 nu)) -- the entries of the reference's Voigt matrix.
__device__ __forceinline__ void le_recipe(const MatParams &m, double E, const double H[3][3], Recipe &r)
{
    const double nu = m.p[1];
    const double f = E / ((1 + nu) * (1 - 2 * nu));
    const double lam = f * nu, mu = f * (1 - 2 * nu) / 2;
    const double e[6] = {H[0][0], H[1][1], H[2][2], 0.5 * (H[0][1] + H[1][0]), 0.5 * (H[0][2] + H[2][0]),
                         0.5 * (H[1][2] + H[2][1])};
    const double tr = e[0] + e[1] + e[2];
    for (int i = 0; i < 3; ++i) r.S[i] = lam * tr + 2 * mu * e[i];
    for (int i = 3; i < 6; ++i) r.S[i] = 2 * mu * e[i];
    r.alpha = lam;
    r.beta = mu;
    r.gam = mu;
    r.c = 0;
}

// J2 radial return with isotropic hardening fy = fy0 + H k + dfy (1 - exp(-delta k)), consistent tangent
// (vonmises.py:186-254). State: plastic strain (engineering Voigt [11,22,33,12,13,23]) + kappa.
// Returns false when the local Newton does not converge in 15 iterations (the reference's CutbackRequest).
__device__ bool j2_recipe(const MatParams &m, double E, const double H[3][3], const double dH[3][3],
                          const double ep[6], double kappa, double ep_t[6], double &kappa_t, Recipe &r)
{
    const double nu = m.p[1], fy0 = m.p[2], HL = m.p[3], dfy = m.p[4], delta = m.p[5];
    const double f = E / ((1 + nu) * (1 - 2 * nu));
    const double lam = f * nu, mu = f * (1 - 2 * nu) / 2;
    const double G = E / (2 * (1.0 + nu));
    // total and incremental strain (tensor components)
    const double e[6] = {H[0][0], H[1][1], H[2][2], 0.5 * (H[0][1] + H[1][0]), 0.5 * (H[0][2] + H[2][0]),
                         0.5 * (H[1][2] + H[2][1])};
    const double de[6] = {dH[0][0], dH[1][1], dH[2][2], 0.5 * (dH[0][1] + dH[1][0]), 0.5 * (dH[0][2] + dH[2][0]),
                          0.5 * (dH[1][2] + dH[2][1])};
    const double epT[6] = {ep[0], ep[1], ep[2], 0.5 * ep[3], 0.5 * ep[4], 0.5 * ep[5]};
    for (int i = 0; i < 6; ++i) ep_t[i] = ep[i];
    for (int i = 0; i < 6; ++i) r.N[i] = 0.0;  // flow direction only on plastic points
    kappa_t = kappa;
    r.alpha = lam;
    r.beta = mu;
    r.gam = mu;
    r.c = 0;
    // |dE| of the engineering-Voigt increment (vonmises.py:211)
    const double dn = sqrt(de[0] * de[0] + de[1] * de[1] + de[2] * de[2] +
                           4 * (de[3] * de[3] + de[4] * de[4] + de[5] * de[5]));
    double ee[6];
    if (dn < 1e-14) {  // stress stays at the converged state, elastic tangent
        for (int i = 0; i < 6; ++i) ee[i] = (e[i] - de[i]) - epT[i];
    } else {
        for (int i = 0; i < 6; ++i) ee[i] = e[i] - epT[i];
    }
    const double tr = ee[0] + ee[1] + ee[2];
    double s[6];
    for (int i = 0; i < 3; ++i) r.S[i] = lam * tr + 2 * mu * ee[i];
    for (int i = 3; i < 6; ++i) r.S[i] = 2 * mu * ee[i];
    if (dn < 1e-14) return true;
    const double p3 = (r.S[0] + r.S[1] + r.S[2]) / 3.0;
    for (int i = 0; i < 3; ++i) s[i] = r.S[i] - p3;
    for (int i = 3; i < 6; ++i) s[i] = r.S[i];
    const double devNorm = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
    const double sq23 = sqrt(2.0 / 3.0), sq6 = sqrt(6.0);
    // the exponential saturation term; identically 0 (and its slope 0) without nonlinear hardening
    const bool sat = (dfy != 0.0) && (delta != 0.0);
    auto fy = [&](double k) { return fy0 + HL * k + (sat ? dfy * (1.0 - exp(-delta * k)) : 0.0); };
    auto dfy_dk = [&](double k) { return HL + (sat ? dfy * delta * exp(-delta * k) : 0.0); };
    if (!(devNorm - sq23 * fy(kappa) > 0.0)) return true;
    double dK = 0;
    int counter = 0;
    for (;;) {
        const double k = kappa + dK;
        const double R = devNorm - sq6 * G * dK - sq23 * fy(k);
        if (!(fabs(R) > 1e-12)) break;
        if (counter == 15) return false;
        const double dR = -sq6 * G - sq23 * dfy_dk(k);
        dK -= R / dR;
        counter++;
    }
    const double dLam = sqrt(1.5) * dK;
    kappa_t = kappa + dK;
    const double inv_norm = 1.0 / devNorm;
    for (int i = 0; i < 6; ++i) r.N[i] = s[i] * inv_norm;
    for (int i = 0; i < 6; ++i) r.S[i] -= 2.0 * G * dLam * r.N[i];
    for (int i = 0; i < 3; ++i) ep_t[i] = ep[i] + dLam * r.N[i];
    for (int i = 3; i < 6; ++i) ep_t[i] = ep[i] + 2.0 * dLam * r.N[i];
    const double hp = dfy_dk(kappa_t);
    const double theta = 2.0 * G * dLam / devNorm;
    const double a1 = 2.0 * G * (1.0 / (1.0 + hp / (3.0 * G)) - theta);
    const double a2 = 2.0 * G * theta;
    r.alpha = lam + a2 / 3.0;
    r.beta = mu - 0.5 * a2;
    r.gam = r.beta;
    r.c = a1;
    return true;
}

// Neo-Hooke Wa (Pence-Gou): tau = mu B + lamBar I, lamBar = (K - 2mu/3)(J^2 - J) - mu,
// muBar = (K - 2mu/3)(2J^2 - J) (neohookepencegouformulationa.py:130-141). With p = dN/dx, r = dN/dX the
// total-Lagrangian tangent of element.py:405-427 reduces to alpha = muBar, beta = -lamBar, gamma = mu.
__device__ __forceinline__ void nh_recipe(double mu, double Kb, const double F[3][3], double J, Recipe &r)
{
    const double lamBar = (Kb - 2.0 / 3.0 * mu) * (J * J - J) - mu;
    const double muBar = (Kb - 2.0 / 3.0 * mu) * (2 * J * J - J);
    double Bc[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bc[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
    r.S[0] = mu * Bc[0][0] + lamBar;
    r.S[1] = mu * Bc[1][1] + lamBar;
    r.S[2] = mu * Bc[2][2] + lamBar;
    r.S[3] = mu * Bc[0][1];
    r.S[4] = mu * Bc[0][2];
    r.S[5] = mu * Bc[1][2];
    r.alpha = muBar;
    r.beta = -lamBar;
    r.gam = mu;
    r.c = 0;
}

// St. Venant-Kirchhoff: S = lambda tr(E) I + 2 mu E (E Green strain), tau = F S F^T. With p = dN/dx, r = dN/dX
// and q = B p (= F r) the total-Lagrangian tangent of element.py:405-427 (dtau/dF of this law) reduces to
//   K_ab,ik = lambda q_a,i q_b,k + mu q_a,k q_b,i + mu (r_a . r_b) B_ik + delta_ik (p_a . tau p_b),
// so the recipe carries alpha = lambda, beta = mu, B in the flow-direction slot and tau as the stress.
__device__ __forceinline__ void svk_recipe(double E, double nu, const double F[3][3], Recipe &r)
{
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    double Eg[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Eg[i][j] = 0.5 * (F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j] - (i == j ? 1.0 : 0.0));
    const double tr = Eg[0][0] + Eg[1][1] + Eg[2][2];
    double S2[3][3], FS[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S2[i][j] = 2 * mu * Eg[i][j] + (i == j ? lam * tr : 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) FS[i][j] = F[i][0] * S2[0][j] + F[i][1] * S2[1][j] + F[i][2] * S2[2][j];
    auto tau = [&](int i, int j) { return FS[i][0] * F[j][0] + FS[i][1] * F[j][1] + FS[i][2] * F[j][2]; };
    auto bc = [&](int i, int j) { return F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2]; };
    r.S[0] = tau(0, 0);
    r.S[1] = tau(1, 1);
    r.S[2] = tau(2, 2);
    r.S[3] = tau(0, 1);
    r.S[4] = tau(0, 2);
    r.S[5] = tau(1, 2);
    r.N[0] = bc(0, 0);
    r.N[1] = bc(1, 1);
    r.N[2] = bc(2, 2);
    r.N[3] = bc(0, 1);
    r.N[4] = bc(0, 2);
    r.N[5] = bc(1, 2);
    r.alpha = lam;
    r.beta = mu;
    r.gam = 0;
    r.c = 0;
}

// Workgroup barrier that orders LDS accesses only. The element kernels share data between their lanes through LDS
// alone; __syncthreads() also orders global memory, i.e. waits until every store issued before it has completed
// (s_waitcnt vmcnt(0)), which would hold the record staging behind the previous stores.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------------------------------------------
// kernel arguments
// ------------------------------------------------------------------------------------------------
struct EvalArgs {
    const int32_t *edof;    // [npe][n_el]
    const int32_t *colpos;  // [n_el][npe*npe] (plan position major: one element's offsets are contiguous)
    const int32_t *order;   // [n_el]
    const int32_t *indptr;  // [n_dof+1]
    const double *el_param;  // [n_el] original element order, or null
    const double *mtab;      // per-element material sets [n_mat][8], or null (then mat holds the parameters)
    const int32_t *mid;      // [n_el] set of the element at each plan position (with mtab)
    const double *X, *U, *dU;
    const double *state_n;
    double *state_t;
    double *P, *F;
    void *vals;
    double *d64;  // EWFE_F32_D64: FP64 corrections of the diagonal blocks (9 per row node), else null
    double *pnewdt;
    int64_t n_el;  // SoA stride
    int64_t p_begin, p_end;
    MatParams mat;
};

// parameters of the element at plan position pos: its material's set when the plan holds a table
__device__ __forceinline__ MatParams mat_at(const EvalArgs &A, int64_t pos)
{
    MatParams m = A.mat;
    if (A.mtab) {
        const double *t = A.mtab + 8 * (int64_t)A.mid[pos];
        for (int i = 0; i < 8; ++i) m.p[i] = t[i];
    }
    return m;
}

__device__ __forceinline__ void atomic_min_pos(double *addr, double v)
{
    // positive doubles order like their bit patterns
    atomicMin(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

template <typename VT>
__device__ __forceinline__ void add_to(VT *vals, int64_t idx, double v)
{
    vals[idx] = vals[idx] + (VT)v;
}

// Phase 1 of the 8-node kernel, one quadrature point: the kinematics from the nodes staged in LDS (accessor
// nd(f, a) gives field f = X0..2, U0..2, dU0..2 of node a; dN(a, i) the parametric shape derivatives at the
// point) and the material update. Returns the inverse Jacobian
// Ji (the gradients are g_a = Ji dN_a), for the total-Lagrangian form also F^-1, wd = w detJ and the recipe.
// Nothing is written to LDS here: the caller stores the gradients once every lane has read the staged nodes.
template <int NN, int FORM, int MAT, typename DN, typename ND>
__device__ __forceinline__ bool hex8_point(DN dN, ND nd, const MatParams &M, double Epar, const double *state_loaded, double *state_held,
                                           double Ji[3][3], double Fi[3][3], Recipe &r, double &wd)
{
    double ep[6] = {0, 0, 0, 0, 0, 0}, kn = 0.0;
    if (FORM == FORM_SS && MAT == M_J2) {
        for (int s = 0; s < 6; ++s) ep[s] = state_loaded[s];
        kn = state_loaded[6];
    }
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < NN; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += dN(a, i) * nd(j, a);
    wd = inv3(J, Ji);  // Gauss weight 1
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dH[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < NN; ++a) {
        double g[3];
        for (int j = 0; j < 3; ++j) g[j] = Ji[j][0] * dN(a, 0) + Ji[j][1] * dN(a, 1) + Ji[j][2] * dN(a, 2);
        for (int i = 0; i < 3; ++i) {
            const double u = nd(3 + i, a);
            const double du = nd(6 + i, a);
            for (int j = 0; j < 3; ++j) {
                H[i][j] += u * g[j];
                if (MAT == M_J2) dH[i][j] += du * g[j];
            }
        }
    }
    bool ok = true;
    if (FORM == FORM_SS) {
        if (MAT == M_LE) {
            le_recipe(M, Epar, H, r);
        } else {
            double ep_t[6], kt;
            ok = j2_recipe(M, Epar, H, dH, ep, kn, ep_t, kt, r);
            for (int s = 0; s < 6; ++s) state_held[s] = ep_t[s];
            state_held[6] = kt;
        }
    } else {
        double F[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
        const double Jd = inv3(F, Fi);
        if (MAT == M_SVK)
            svk_recipe(Epar, M.p[1], F, r);
        else
            nh_recipe(Epar, M.p[1], F, Jd, r);
    }
    return ok;
}

// Phase 1 of the 20-node kernel with four lanes per quadrature point: lane s of the point's group handles nodes s,
// s + 4, ..., the node sums (J, grad u, grad du) are completed by a fixed butterfly over the group, so all four lanes
// hold bitwise the same sums and evaluate the material identically; lane s = 0 alone writes the point's state and
// reports a cutback. The group's lanes must be active together (the butterfly reads all four).
template <int NN, int FORM, int MAT, typename ND, typename PG, typename PP>
__device__ __forceinline__ bool qp_phase_coop(const double (*dN)[3], double w, int s, ND nd, PG put_g, PP put_p,
                                              const EvalArgs &A, const MatParams &M, double Epar, int64_t pos, int q,
                                              int NQ, Recipe &r, double &wd)
{
    constexpr int NS = NN / 4;  // nodes per lane
    double ep[6] = {0, 0, 0, 0, 0, 0}, kn = 0.0;
    if (FORM == FORM_SS && MAT == M_J2) {
        const int64_t n = A.n_el;
        for (int c = 0; c < 6; ++c) ep[c] = A.state_n[((int64_t)c * NQ + q) * n + pos];
        kn = A.state_n[((int64_t)6 * NQ + q) * n + pos];
    }
    auto group_sum = [](double v) {
        v += __shfl_xor(v, 1, 4);
        v += __shfl_xor(v, 2, 4);
        return v;
    };
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ji[3][3];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int a = s + 4 * t;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * nd(j, a);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] = group_sum(J[i][j]);
    const double detJ = inv3(J, Ji);
    wd = w * detJ;
    double g[NS][3];
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dH[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int a = s + 4 * t;
        for (int j = 0; j < 3; ++j) g[t][j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
        for (int j = 0; j < 3; ++j) put_g(a, j, g[t][j]);
        for (int i = 0; i < 3; ++i) {
            const double u = nd(3 + i, a);
            const double du = (MAT == M_J2) ? nd(6 + i, a) : 0.0;
            for (int j = 0; j < 3; ++j) {
                H[i][j] += u * g[t][j];
                if (MAT == M_J2) dH[i][j] += du * g[t][j];
            }
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            H[i][j] = group_sum(H[i][j]);
            if (MAT == M_J2) dH[i][j] = group_sum(dH[i][j]);
        }
    bool ok = true;
    if (FORM == FORM_SS) {
        if (MAT == M_LE) {
            le_recipe(M, Epar, H, r);
        } else {
            double ep_t[6], kt;
            const int64_t n = A.n_el;
            ok = j2_recipe(M, Epar, H, dH, ep, kn, ep_t, kt, r);
            if (s == 0) {
                for (int c = 0; c < 6; ++c) A.state_t[((int64_t)c * NQ + q) * n + pos] = ep_t[c];
                A.state_t[((int64_t)6 * NQ + q) * n + pos] = kt;
            }
        }
    } else {
        double F[3][3], Fi[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
        const double Jd = inv3(F, Fi);
        if (MAT == M_SVK)
            svk_recipe(Epar, M.p[1], F, r);
        else
            nh_recipe(Epar, M.p[1], F, Jd, r);
        // spatial gradients NA_a = F^-T g_a of the lane's nodes
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const int a = s + 4 * t;
            for (int m = 0; m < 3; ++m) put_p(a, m, g[t][0] * Fi[0][m] + g[t][1] * Fi[1][m] + g[t][2] * Fi[2][m]);
        }
    }
    return ok;
}

// Number of LDS fields per quadrature point: g (3 NN) [+ p (3 NN) for TL] + alpha,beta,gam,c + N (6) + S (6)
template <int NN, int FORM>
struct QPLayout {
    static constexpr int G = 0;
    static constexpr int P = (FORM == FORM_TL) ? 3 * NN : 0;  // offset of p (aliases g for small strain)
    static constexpr int COEF = (FORM == FORM_TL) ? 6 * NN : 3 * NN;
    static constexpr int NV = COEF + 4;
    static constexpr int S = NV + 6;
    static constexpr int NF = S + 6;
};

// ------------------------------------------------------------------------------------------------
// C3D8 / C3D8TL, pass 1: element records. 8 elements per 64-lane workgroup.
// Record of element e: kbuf[e][300] = the 36 unique 3x3 blocks K_ab (a <= b) in pair order
// pair8(a, b) = a 8 - a (a + 1) / 2 + b; off-diagonal blocks row-major (9 values), diagonal blocks (symmetric)
// as their upper triangle (00, 01, 02, 11, 12, 22: 6 values); fe[e][8][3] = internal nodal forces f_a
// (P_e = -f).
// ------------------------------------------------------------------------------------------------
constexpr int kPairs8 = 36;
constexpr int kRec8 = 28 * 9 + 8 * 6;  // 300

__host__ __device__ constexpr int pair8(int lo, int hi) { return lo * 8 - lo * (lo + 1) / 2 + hi; }
// offset of block (lo, hi), lo <= hi, in a record: the diagonal blocks before it are (i, i) for i < lo, and
// (lo, lo) itself when hi > lo
__host__ __device__ constexpr int block_off8(int lo, int hi) { return 9 * pair8(lo, hi) - 3 * (lo + (hi > lo ? 1 : 0)); }
constexpr int kSym6[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};  // (i, k) -> upper-triangle slot

// Records of 20-node elements, the same layout: the 210 unique blocks K_ab (a <= b) in pair order
// pair20(a, b) = a 20 - a (a + 1) / 2 + b, 190 off-diagonal blocks of 9 values and 20 diagonal upper triangles of
// 6: 1830 values, padded to 1832 so that FP32 records are whole 16-byte units too.
constexpr int kRec20 = 190 * 9 + 20 * 6 + 2;  // 1832
// values after the last 20-node record of the buffer: the 9 values loaded for the last diagonal block (19, 19) run
// one value past its record
constexpr int kRecPad20 = 2;
__host__ __device__ constexpr int pair20(int lo, int hi) { return lo * 20 - lo * (lo + 1) / 2 + hi; }
__host__ __device__ constexpr int block_off20(int lo, int hi) { return 9 * pair20(lo, hi) - 3 * (lo + (hi > lo ? 1 : 0)); }

// Reads block K_ab of a record oriented for row node a (transposed when the record holds K_ba). Every lane
// loads 9 values, whatever the block kind, so the loads of a wave do not diverge; for a diagonal block the
// last 3 belong to the next block (or to the padding after the last record) and are not used.
constexpr int kRecPad8 = 3;  // values after the last record of the buffer
template <typename VT>
__device__ __forceinline__ void load_block8(const VT *rec, int a, int b, double v[9])
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const VT *src = rec + block_off8(lo, hi);
    double u[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) u[t] = (double)src[t];
    const bool diag = a == b, tr = a > b;  // tr: the stored block is K_ba, its transpose is K_ab
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) v[3 * i + k] = diag ? u[kSym6[i][k]] : (tr ? u[3 * k + i] : u[3 * i + k]);
}

// block K_ab of a 20-node record oriented for row node a (load_block8's convention; the last 3 values loaded for a
// diagonal block belong to the next block or to the padding and are not used)
template <typename VT>
__device__ __forceinline__ void load_block20(const VT *rec, int a, int b, double v[9])
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const VT *src = rec + block_off20(lo, hi);
    double u[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) u[t] = (double)src[t];
    const bool diag = a == b, tr = a > b;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) v[3 * i + k] = diag ? u[kSym6[i][k]] : (tr ? u[3 * k + i] : u[3 * i + k]);
}

struct RecordArgs {
    void *kbuf;
    double *fe;
    // this launch evaluates the plan positions [q_begin, q_end), 8 per workgroup; the batches are dealt to
    // the 8 XCDs in contiguous runs of xcd_batches (workgroups go to XCDs round robin)
    int64_t q_begin, q_end, xcd_batches;
};

// LDS layout of the per-quadrature-point data of k_hex8_elem, per field f: 64 values, one per (element le,
// quadrature point q), in the precision of the tangent arithmetic (doubles, or floats for the FP32 tangent).
// Fields: per-node gradients g (and p for the total-Lagrangian form), 3 per node, then the tangent
// coefficients alpha wd, beta wd [, gamma wd for the total-Lagrangian form; gamma = beta for small strain]
// [, c wd and the flow direction N for J2] [, the stress S wd for St. Venant-Kirchhoff, whose tangent has a
// stress term] [, XF fields of staging space for the matrix-core tangent, from field XS].
template <int FORM, int MAT, int XF = 0>
struct Hex8Layout {
    // HAS_N: the per-point 6-vector slot NV (the J2 flow direction, or B = F F^T for St. Venant-Kirchhoff)
    static constexpr bool TL = FORM == FORM_TL, HAS_N = MAT == M_J2 || MAT == M_SVK;
    static constexpr bool HAS_S = MAT == M_SVK;
    static constexpr int G = 0;
    static constexpr int P = TL ? 24 : 0;  // aliases g for small strain
    static constexpr int AL = TL ? 48 : 24;
    static constexpr int BE = AL + 1;
    static constexpr int GA = TL ? AL + 2 : BE;
    static constexpr int CC = TL ? AL + 3 : AL + 2;  // only with HAS_N
    static constexpr int NV = CC + 1;                // 6, only with HAS_N
    static constexpr int S = HAS_N ? NV + 6 : CC;    // 6, only with HAS_S
    static constexpr int XS = HAS_S ? S + 6 : S;
    static constexpr int NF = XS + XF;
};

// staging fields of the matrix-core tangent: T + G and Tb in record slot order (2 x 300 floats) and S_ab (36 floats)
template <typename PT>
constexpr int mfma_staging_fields()
{
    return (int)((2 * 300 + 36) * sizeof(float) + 64 * sizeof(PT) - 1) / (int)(64 * sizeof(PT));
}

// A 64-lane ds_*_b64 is served 16 lanes at a time by 32 four-byte banks, i.e. 16 double slots. Phase 1 writes
// a field with lanes (le, q) (fixed field), phase 2 reads node fields with lanes (le, a) at a fixed q, where
// lane a reads node b = (a + j) mod 8: a 16-lane group (le in {2g, 2g + 1}) then touches all 8 nodes b. Node
// fields are XOR-swizzled by 2 b inside their 16-slot group so that both access patterns hit 16 different
// slots; scalar fields are read as broadcasts (one address per element) and need no swizzle. With float fields
// (ds_*_b32, 32 lanes at a time over the 32 banks) the same indices are conflict-free too: lanes 0..31 are
// (le, node) = (0..3, 0..7) and land on 16 (le >> 1) + 2 (q ^ b) + (le & 1), 32 different banks.
__device__ __forceinline__ int slot_scalar(int f, int q, int le) { return f * 64 + 16 * (le >> 1) + 2 * q + (le & 1); }
__device__ __forceinline__ int slot_node(int base, int b, int i, int q, int le)
{
    return (base + 3 * b + i) * 64 + 16 * (le >> 1) + ((2 * q + (le & 1)) ^ (2 * b));
}

// Phase 2 of k_hex8_elem for FP32 records computed on the matrix cores (tangent mode EWFE_F32_MFMA).
// With the per-point fields of phase 1 in LDS, K_e = T + B' + G, where over the quadrature points q
//   T_mn  = sum_q (alpha_q p_m p_n - c_q v_m v_n)      (m = 3a + i, n = 3b + k: the 24 element dofs)
//   B'_mn = sum_q beta_q p_(3a+k) p_(3b+i)               (the 3x3 blocks of Tb = sum_q beta_q p p^T, transposed)
//   G_mn  = delta_ik S_ab,  S_ab = sum_q gamma_q (r_a . r_b)
// T and Tb are symmetric rank-8 (rank-16 for J2) products, formed by v_mfma_f32_16x16x4_f32 on the three 16x16
// tiles (0,0), (0,1), (1,1) that cover the upper triangle of the 24x24 matrix (rows and columns 24..31 are
// zero), lane quarter g supplying points q = g and g + 4; S is one more 16x16 tile (rows 8..15 zero) with the 24
// products (q, j) as its inner dimension: 24 MFMAs for J2, 18 otherwise, per element. The nodal forces fa come
// from phase 1 in FP64, exactly as in the other variants (P and F are bitwise the same). T + G and Tb are staged
// in LDS in the record's slot order, each slot written by one lane, and summed when the record is written out
// (deterministic). One element at a time; staging fields at the end of the layout (not written by phase 1) hold the
// staging. PT: the per-point data's precision (floats: the operands need no rounding).
template <int FORM, int MAT, typename VT, typename PT>
__device__ __forceinline__ void hex8_tangent_mfma(const EvalArgs &A, const RecordArgs &R, PT *s_qp,
                                                  int32_t *s_el, int le, int t, bool active, int64_t q0, int64_t el,
                                                  const double (&fa)[3], const double (&st_t)[7])
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int NQ = 8, EPB = 8;
    using L = Hex8Layout<FORM, MAT, mfma_staging_fields<PT>()>;
    constexpr int PB = L::TL ? L::P : L::G;  // the gradient the tangent's p terms use
    const int lane = threadIdx.x;
    const int a = t;
    const int64_t n_el = A.n_el;
    const int64_t pos = q0 + le;
    auto ld = [&](int f, int q, int e) { return s_qp[slot_scalar(f, q, e)]; };
    auto ldg = [&](int base, int b, int i, int q, int e) { return s_qp[slot_node(base, b, i, q, e)]; };

    if (t == 0) s_el[le] = (int32_t)el;
    if (active)
        for (int i = 0; i < 3; ++i) R.fe[el * 24 + 3 * a + i] = fa[i];
    if (active && FORM == FORM_SS && MAT == M_J2)
        for (int s = 0; s < 7; ++s) A.state_t[((int64_t)s * NQ + t) * n_el + pos] = st_t[s];
    lds_barrier();  // s_el is written
    float *s_ta = reinterpret_cast<float *>(s_qp + L::XS * 64);  // T + G in record slot order (300 values)
    float *s_tb = s_ta + kRec8;                                    // Tb in record slot order
    float *s_sg = s_tb + kRec8;                                    // S_ab of the pairs a <= b (36 values)
    static_assert((2 * kRec8 + kPairs8) * sizeof(float) <= mfma_staging_fields<PT>() * 64 * sizeof(PT),
                  "the staging must fit in its fields");
    static_assert((kRec8 * 4) % 16 == 0, "staged records must be 16-byte aligned");

    // (B) one element at a time: lane (g, c) supplies rows c and 16 + c for points g and g + 4
    const int g = lane >> 4, c = lane & 15;
    const int m0 = c, m1 = 16 + c;
    const bool has1 = m1 < 24;
    const int a0 = m0 / 3, i0 = m0 % 3, a1 = has1 ? m1 / 3 : 0, i1 = has1 ? m1 % 3 : 0;
    const int nvalid = (R.q_end - q0) < EPB ? (int)(R.q_end - q0) : EPB;
    struct Ops {
        float x0[2], x1[2], al[2], be[2], y0[2], y1[2], mc[2];
    };
    auto operands = [&](int e, Ops &o) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = g + 4 * h;
            o.x0[h] = (float)ldg(PB, a0, i0, q, e);
            o.x1[h] = has1 ? (float)ldg(PB, a1, i1, q, e) : 0.f;
            o.al[h] = (float)ld(L::AL, q, e);
            o.be[h] = (float)ld(L::BE, q, e);
            o.y0[h] = o.y1[h] = o.mc[h] = 0.f;
            if (L::HAS_N) {
                // v = N p with the flow direction N (tensor components [11,22,33,12,13,23])
                float Nv[6];
                for (int k = 0; k < 6; ++k) Nv[k] = (float)ld(L::NV + k, q, e);
                auto vrow = [&](int aa, int ii) {
                    const int j0 = ii == 0 ? 0 : (ii == 1 ? 3 : 4), j1 = ii == 0 ? 3 : (ii == 1 ? 1 : 5),
                              j2 = ii == 0 ? 4 : (ii == 1 ? 5 : 2);
                    return Nv[j0] * (float)ldg(PB, aa, 0, q, e) + Nv[j1] * (float)ldg(PB, aa, 1, q, e) +
                           Nv[j2] * (float)ldg(PB, aa, 2, q, e);
                };
                o.y0[h] = vrow(a0, i0);
                o.y1[h] = has1 ? vrow(a1, i1) : 0.f;
                o.mc[h] = -(float)ld(L::CC, q, e);
            }
        }
    };
    Ops cur;
    if (nvalid > 0) operands(0, cur);
    for (int e = 0; e < nvalid; ++e) {
        f4 acc[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) acc[u] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float x0 = cur.x0[h], x1 = cur.x1[h], al = cur.al[h], be = cur.be[h];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(al * x0, x0, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(al * x0, x1, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(al * x1, x1, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(be * x0, x0, acc[3], 0, 0, 0);
            acc[4] = __builtin_amdgcn_mfma_f32_16x16x4f32(be * x0, x1, acc[4], 0, 0, 0);
            acc[5] = __builtin_amdgcn_mfma_f32_16x16x4f32(be * x1, x1, acc[5], 0, 0, 0);
            if (L::HAS_N) {
                const float y0 = cur.y0[h], y1 = cur.y1[h], mc = cur.mc[h];
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(mc * y0, y0, acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(mc * y0, y1, acc[1], 0, 0, 0);
                acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(mc * y1, y1, acc[2], 0, 0, 0);
            }
        }
        // S: rows a = c (< 8), inner index k = 4 s + g over (q, j) = (k / 3, k % 3)
#pragma unroll
        for (int st = 0; st < 6; ++st) {
            const int k = 4 * st + g, q = k / 3, j = k % 3;
            const float r = c < 8 ? (float)ldg(L::G, c, j, q, e) : 0.f;
            const float gq = (float)ld(L::GA, q, e);
            acc[6] = __builtin_amdgcn_mfma_f32_16x16x4f32(gq * r, r, acc[6], 0, 0, 0);
        }
        // the next element's operands, loaded while this element's products drain
        Ops nxt = cur;
        if (e + 1 < nvalid) operands(e + 1, nxt);
        // S_ab: lane (g, c) holds rows a = 4 g + v, column b = c
        if (g < 2 && c < 8)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ra = 4 * g + v;
                if (ra <= c) s_sg[pair8(ra, c)] = acc[6][v];
            }
        lds_barrier();
        // accumulator u % 3 = tile (0,0), (0,1), (1,1); lane (g, c) holds rows 16 R + 4 g + v, column 16 C + c.
        // Values with m <= n < 24 are the record's: block (a, b) = (m / 3, n / 3), a <= b. T + G goes to the slot of
        // K_ab[i][k] (i = m % 3, k = n % 3); the Tb value (m, n) = ((a, x), (b, y)) belongs to K_ab[y][x].
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int n = (u == 0 ? 0 : 16) + c;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = (u == 2 ? 16 : 0) + 4 * g + v;
                if (m <= n && n < 24) {
                    const int am = m / 3, i = m % 3, bn = n / 3, k = n % 3;
                    float val = acc[u][v];
                    if (i == k) val += s_sg[pair8(am, bn)];
                    const bool offd = am < bn;
                    s_ta[offd ? block_off8(am, bn) + 3 * i + k : block_off8(am, am) + kSym6[i][k]] = val;
                    s_tb[offd ? block_off8(am, bn) + 3 * k + i : block_off8(am, am) + kSym6[i][k]] = acc[3 + u][v];
                }
            }
        }
        lds_barrier();
        // the record = staged T + G plus staged Tb: 75 16-byte units, non-temporal (read back by the gather pass)
        constexpr int NW = kRec8 / 4;
        static_assert(kRec8 % 4 == 0, "records are whole 16-byte units");
        VT *dst = reinterpret_cast<VT *>(R.kbuf) + (int64_t)s_el[e] * kRec8;
        for (int w = lane; w < NW; w += 64) {
            const f4 sum = reinterpret_cast<const f4 *>(s_ta)[w] + reinterpret_cast<const f4 *>(s_tb)[w];
            __builtin_nontemporal_store(sum, reinterpret_cast<f4 *>(dst) + w);
        }
        lds_barrier();  // the staging is free for the next element
        cur = nxt;
    }
}

// Precision of the per-point data and of the tangent arithmetic of k_hex8_elem: that of the records, also for the
// matrix-core tangent, whose operands are FP32 (the float fields need no rounding when they are packed). (The J2
// consistent tangent has the deviatoric stiffness 2G - a2 - a1, about 2H/3, along the flow direction as a difference
// of terms of size 2G, so FP32 arithmetic puts ~240 times the rounding of its terms into that soft mode; with the
// assembled FP32 diagonal blocks the 128^3 J2 schedule needed 12 instead of 7 solves, with the FP64 diagonal
// corrections of EWFE_F32_D64, which make every row sum to zero again, it needs the FP64 tangent's 7.)
template <int MAT, typename VT, bool MF>
using hex8_point_t = VT;

// occupancy the compiler is asked for: float point data and accumulators fit 3 waves per SIMD (4 waves need
// register spills)
template <int MAT, typename VT, bool MF>
constexpr int hex8_min_waves()
{
    return sizeof(hex8_point_t<MAT, VT, MF>) == 4 ? 3 : 1;
}

// MT: the plan holds a material table (a separate instantiation, so that the single-material path carries no
// per-element parameter selects)
// MF: FP32 records with the tangent formed on the matrix cores (only with VT = float)
// PT: the precision of the per-point data in LDS and of the tangent arithmetic of phase 2 (hex8_point_t)
template <int FORM, int MAT, typename VT, bool MT, bool MF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(hex8_min_waves<MAT, VT, MF>()))) void k_hex8_elem(
    EvalArgs A, RecordArgs R)
{
    constexpr int NN = 8, NQ = 8, EPB = 8;
    using PT = hex8_point_t<MAT, VT, MF>;
    using L = Hex8Layout<FORM, MAT, MF ? mfma_staging_fields<PT>() : 0>;
    constexpr int QP_WORDS = L::NF * 64;
    static_assert((EPB / 2) * kRec8 * sizeof(VT) <= QP_WORDS * sizeof(PT),
                  "record staging must fit in the quadrature-point area");
    __shared__ __align__(16) PT s_qp[QP_WORDS];
    __shared__ int32_t s_el[EPB];
    const int lane = threadIdx.x;
    const int le = lane >> 3, t = lane & 7;
    const int64_t xb = blockIdx.x >> 3;
    if (xb >= R.xcd_batches) return;  // uniform over the workgroup
    const int64_t q0 = R.q_begin + ((int64_t)(blockIdx.x & 7) * R.xcd_batches + xb) * EPB;
    if (q0 >= R.q_end) return;  // uniform over the workgroup
    const int64_t pos = q0 + le;
    const bool active = pos < R.q_end;
    const int64_t n = A.n_el;
    const int64_t el = active ? (int64_t)A.order[pos] : 0;  // records and nodal forces are stored by element

    // phase 1a: lane (le, t) stages node t of its element (X, U (, dU)) in LDS, over the gradient fields (they are
    // written only in phase 1b); then lane (le, q = t) evaluates one quadrature point reading the element's nodes
    // from there. One node per lane keeps the nodal data out of the registers.
    constexpr int NVN = (MAT == M_J2) ? 9 : 6;  // staged values per node
    static_assert(EPB * NN * NVN * sizeof(double) <= 24 * 64 * sizeof(PT), "nodal staging must fit in the gradients");
    double *s_nd = reinterpret_cast<double *>(s_qp);
    const MatParams M = MT ? mat_at(A, active ? pos : q0) : A.mat;
    double Epar = M.p[0];
    if (active && A.el_param) Epar = A.el_param[A.order[pos]];
    double st[7] = {0, 0, 0, 0, 0, 0, 0};  // converged J2 state of point (le, q = t), loaded with the nodes
    // its trial state, stored after phase 2: stored at once, the next wait for a load (which on this GPU also
    // waits for every earlier store) would hold phase 2 until the stores are done
    double st_t[7] = {0, 0, 0, 0, 0, 0, 0};
    if (active && FORM == FORM_SS && MAT == M_J2)
        for (int s = 0; s < 7; ++s) st[s] = A.state_n[((int64_t)s * NQ + t) * n + pos];
    if (active) {
        const int32_t da = A.edof[(int64_t)t * n + pos];
        double *dst = s_nd + (le * NN + t) * NVN;
        const double x0 = A.X[da], x1 = A.X[da + 1], x2 = A.X[da + 2];
        const double u0 = A.U[da], u1 = A.U[da + 1], u2 = A.U[da + 2];
        double d0 = 0, d1 = 0, d2 = 0;
        if (MAT == M_J2) {
            d0 = A.dU[da];
            d1 = A.dU[da + 1];
            d2 = A.dU[da + 2];
        }
        dst[0] = x0;
        dst[1] = x1;
        dst[2] = x2;
        dst[3] = u0;
        dst[4] = u1;
        dst[5] = u2;
        if (MAT == M_J2) {
            dst[6] = d0;
            dst[7] = d1;
            dst[8] = d2;
        }
    }
    lds_barrier();
    const int q = t;
    // parametric derivatives of the trilinear shape functions at Gauss point q (the points in the order of the
    // host tables; the same arithmetic, so the same values), recomputed where they are used
    const double gp = 1.0 / sqrt(3.0);
    const double xq0 = q < 4 ? -gp : gp, xq1 = (q & 2) ? gp : -gp, xq2 = ((q ^ (q >> 1)) & 1) ? gp : -gp;
    auto dN = [&](int b, int i) {
        const double c0 = ((b & 3) == 1 || (b & 3) == 2) ? 1.0 : -1.0;
        const double c1 = ((b & 3) >= 2) ? 1.0 : -1.0;
        const double c2 = (b >= 4) ? 1.0 : -1.0;
        const double f0 = 1 + c0 * xq0, f1 = 1 + c1 * xq1, f2 = 1 + c2 * xq2;
        return i == 0 ? 0.125 * c0 * f1 * f2 : (i == 1 ? 0.125 * f0 * c1 * f2 : 0.125 * f0 * f1 * c2);
    };
    Recipe r = {};
    double wd = 0.0;
    double Ji[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Fi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (active) {
        const double *src = s_nd + le * NN * NVN;
        auto nd = [&](int f, int a) { return (MAT != M_J2 && f >= 6) ? 0.0 : src[a * NVN + f]; };
        const bool ok = hex8_point<NN, FORM, MAT>(dN, nd, M, Epar, st, st_t, Ji, Fi, r, wd);
        if (!ok) atomic_min_pos(A.pnewdt, 0.5);
    }
    lds_barrier();  // every lane has read the staged nodes

    // phase 1b: lane (le, q) stores its point's gradients and tangent coefficients in LDS, in the tangent's
    // precision, and forms the point's contributions wd S p_a to the nodal forces in FP64. The forces are summed
    // over the element's 8 points by a fixed reduce-scatter over its 8 lanes (lane distance 4, 2, 1), after which
    // lane (le, a) holds f_a; the order is fixed, so P and F are bitwise the same in every precision variant.
    // Inactive elements run it on zeros (the exchanges stay inside an element's lanes).
    double wS[6];
    for (int i = 0; i < 6; ++i) wS[i] = wd * r.S[i];
    auto grad = [&](int a, double c[3]) {
        double g[3], p[3];
        for (int j = 0; j < 3; ++j) g[j] = Ji[j][0] * dN(a, 0) + Ji[j][1] * dN(a, 1) + Ji[j][2] * dN(a, 2);
        for (int j = 0; j < 3; ++j) s_qp[slot_node(L::G, a, j, q, le)] = (PT)g[j];
        if (FORM == FORM_TL) {  // spatial gradients p_a = F^-T g_a
            for (int m = 0; m < 3; ++m) p[m] = g[0] * Fi[0][m] + g[1] * Fi[1][m] + g[2] * Fi[2][m];
            for (int m = 0; m < 3; ++m) s_qp[slot_node(L::P, a, m, q, le)] = (PT)p[m];
        } else {
            for (int m = 0; m < 3; ++m) p[m] = g[m];
        }
        c[0] = wS[0] * p[0] + wS[3] * p[1] + wS[4] * p[2];
        c[1] = wS[3] * p[0] + wS[1] * p[1] + wS[5] * p[2];
        c[2] = wS[4] * p[0] + wS[5] * p[1] + wS[2] * p[2];
    };
    const bool hi4 = (t & 4) != 0, hi2 = (t & 2) != 0, hi1 = (t & 1) != 0;
    double h4[4][3];  // lane distance 4: sums of two points for the nodes (t & 4) | j
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double c0[3], c1[3];
        grad(j, c0);
        grad(j + 4, c1);
        for (int i = 0; i < 3; ++i) {
            const double keep = hi4 ? c1[i] : c0[i], send = hi4 ? c0[i] : c1[i];
            h4[j][i] = keep + __shfl_xor(send, 4, 8);
        }
    }
    double h2[2][3];  // lane distance 2: sums of four points for the nodes (t & 6) | k
#pragma unroll
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 3; ++i) {
            const double keep = hi2 ? h4[2 + k][i] : h4[k][i], send = hi2 ? h4[k][i] : h4[2 + k][i];
            h2[k][i] = keep + __shfl_xor(send, 2, 8);
        }
    double fa[3];  // lane distance 1: node t, all 8 points
    for (int i = 0; i < 3; ++i) {
        const double keep = hi1 ? h2[1][i] : h2[0][i], send = hi1 ? h2[0][i] : h2[1][i];
        fa[i] = keep + __shfl_xor(send, 1, 8);
    }
    {
        auto put = [&](int f, double v) { s_qp[slot_scalar(f, q, le)] = (PT)v; };
        put(L::AL, wd * r.alpha);
        put(L::BE, wd * r.beta);
        if (L::TL) put(L::GA, wd * r.gam);
        if (L::HAS_N) {
            put(L::CC, wd * r.c);
            for (int i = 0; i < 6; ++i) put(L::NV + i, r.N[i]);
        }
        if (MAT == M_SVK)
            for (int i = 0; i < 6; ++i) put(L::S + i, wS[i]);
    }
    lds_barrier();

    if constexpr (MF) {
        static_assert(sizeof(VT) == 4, "the matrix-core tangent writes FP32 records");
        hex8_tangent_mfma<FORM, MAT, VT, PT>(A, R, s_qp, s_el, le, t, active, q0, el, fa, st_t);
        return;
    }
    // phase 2: lane (le, a = t) builds the blocks K_ab, b = a + j (mod 8), j < nb, in the precision PT
    const int a = t;
    const int nb = (a < 4) ? 5 : 4;
    PT K[5][3][3];
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) K[j][i][k] = 0;
    auto ld = [&](int f, int q) { return s_qp[slot_scalar(f, q, le)]; };
    auto ldg = [&](int base, int b, int i, int q) { return s_qp[slot_node(base, b, i, q, le)]; };
    if (active && FORM == FORM_SS && MAT == M_LE) {
        // linear elasticity: alpha = lambda wd, beta = gamma = mu wd with constant lambda, mu, so with the
        // single sum T_ab = sum_q mu wd g_a g_b^T:  K_ab = (lambda / mu) T + T^T + tr(T) I
#pragma unroll 2
        for (int q = 0; q < NQ; ++q) {
            const PT muwd = ld(L::BE, q);
            const PT ga0 = ldg(L::G, a, 0, q), ga1 = ldg(L::G, a, 1, q), ga2 = ldg(L::G, a, 2, q);
            const PT wa[3] = {muwd * ga0, muwd * ga1, muwd * ga2};
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                if (j < nb) {
                    const int b = (a + j) & 7;
                    const PT gb0 = ldg(L::G, b, 0, q), gb1 = ldg(L::G, b, 1, q), gb2 = ldg(L::G, b, 2, q);
                    for (int i = 0; i < 3; ++i) {
                        K[j][i][0] += wa[i] * gb0;
                        K[j][i][1] += wa[i] * gb1;
                        K[j][i][2] += wa[i] * gb2;
                    }
                }
            }
        }
        const double nu = M.p[1];
        const PT ratio = (PT)(2.0 * nu / (1.0 - 2.0 * nu));  // lambda / mu
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            PT T[3][3];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) T[i][k] = K[j][i][k];
            const PT tr = T[0][0] + T[1][1] + T[2][2];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) K[j][i][k] = ratio * T[i][k] + T[k][i] + (i == k ? tr : (PT)0);
        }
    } else if (active && MAT == M_SVK) {
        // St. Venant-Kirchhoff (total Lagrangian):
        // K_ab,ik = alpha q_a,i q_b,k + beta q_a,k q_b,i + beta (r_a . r_b) B_ik + delta_ik (p_a . S p_b),
        // q = B p, with alpha = lambda wd, beta = mu wd, S = wd tau, B unscaled
        for (int q = 0; q < NQ; ++q) {
            const PT al = ld(L::AL, q), be = ld(L::BE, q);
            PT pa[3], ra[3], Bm[3][3], Sm[3][3];
            for (int i = 0; i < 3; ++i) {
                ra[i] = ldg(L::G, a, i, q);
                pa[i] = ldg(L::P, a, i, q);
            }
            {
                const PT b0 = ld(L::NV + 0, q), b1 = ld(L::NV + 1, q), b2 = ld(L::NV + 2, q);
                const PT b3 = ld(L::NV + 3, q), b4 = ld(L::NV + 4, q), b5 = ld(L::NV + 5, q);
                Bm[0][0] = b0, Bm[1][1] = b1, Bm[2][2] = b2;
                Bm[0][1] = Bm[1][0] = b3, Bm[0][2] = Bm[2][0] = b4, Bm[1][2] = Bm[2][1] = b5;
                const PT S0 = ld(L::S + 0, q), S1 = ld(L::S + 1, q), S2 = ld(L::S + 2, q);
                const PT S3 = ld(L::S + 3, q), S4 = ld(L::S + 4, q), S5 = ld(L::S + 5, q);
                Sm[0][0] = S0, Sm[1][1] = S1, Sm[2][2] = S2;
                Sm[0][1] = Sm[1][0] = S3, Sm[0][2] = Sm[2][0] = S4, Sm[1][2] = Sm[2][1] = S5;
            }
            PT qa[3], tpa[3];
            for (int i = 0; i < 3; ++i) {
                qa[i] = Bm[i][0] * pa[0] + Bm[i][1] * pa[1] + Bm[i][2] * pa[2];
                tpa[i] = Sm[i][0] * pa[0] + Sm[i][1] * pa[1] + Sm[i][2] * pa[2];
            }
            {  // diagonal block: symmetric, upper triangle accumulated and mirrored after the loop
                const PT rr = be * (ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
                const PT d1 = pa[0] * tpa[0] + pa[1] * tpa[1] + pa[2] * tpa[2];
                for (int i = 0; i < 3; ++i)
                    for (int k = i; k < 3; ++k) {
                        K[0][i][k] = fma((al + be) * qa[i], qa[k], K[0][i][k]);
                        K[0][i][k] = fma(rr, Bm[i][k], K[0][i][k]);
                    }
                for (int i = 0; i < 3; ++i) K[0][i][i] += d1;
            }
#pragma unroll
            for (int j = 1; j < 5; ++j) {
                if (j < nb) {
                    const int b = (a + j) & 7;
                    PT pb[3], rb[3], qb[3];
                    for (int i = 0; i < 3; ++i) {
                        rb[i] = ldg(L::G, b, i, q);
                        pb[i] = ldg(L::P, b, i, q);
                    }
                    for (int i = 0; i < 3; ++i) qb[i] = Bm[i][0] * pb[0] + Bm[i][1] * pb[1] + Bm[i][2] * pb[2];
                    const PT rr = be * (ra[0] * rb[0] + ra[1] * rb[1] + ra[2] * rb[2]);
                    const PT d1 = tpa[0] * pb[0] + tpa[1] * pb[1] + tpa[2] * pb[2];
                    for (int i = 0; i < 3; ++i)
                        for (int k = 0; k < 3; ++k) {
                            K[j][i][k] = fma(al * qa[i], qb[k], K[j][i][k]);
                            K[j][i][k] = fma(be * qa[k], qb[i], K[j][i][k]);
                            K[j][i][k] = fma(rr, Bm[i][k], K[j][i][k]);
                        }
                    for (int i = 0; i < 3; ++i) K[j][i][i] += d1;
                }
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < i; ++k) K[0][i][k] = K[0][k][i];
    } else if (active) {
        for (int q = 0; q < NQ; ++q) {
            const PT al = ld(L::AL, q), be = ld(L::BE, q), ga = ld(L::GA, q);
            const PT cc = L::HAS_N ? ld(L::CC, q) : (PT)0;
            PT pa[3], ra[3], va[3] = {0, 0, 0}, Nq[6];
            for (int i = 0; i < 3; ++i) {
                ra[i] = ldg(L::G, a, i, q);
                pa[i] = L::TL ? ldg(L::P, a, i, q) : ra[i];
            }
            if (L::HAS_N) {
                for (int i = 0; i < 6; ++i) Nq[i] = ld(L::NV + i, q);
                va[0] = Nq[0] * pa[0] + Nq[3] * pa[1] + Nq[4] * pa[2];
                va[1] = Nq[3] * pa[0] + Nq[1] * pa[1] + Nq[5] * pa[2];
                va[2] = Nq[4] * pa[0] + Nq[5] * pa[1] + Nq[2] * pa[2];
            }
            PT Aa[3], Ba[3], Va[3];
            for (int i = 0; i < 3; ++i) {
                Aa[i] = al * pa[i];
                Ba[i] = be * pa[i];
                Va[i] = cc * va[i];
            }
            {  // diagonal block K_aa = (alpha + beta) p_a p_a^T + gamma |r_a|^2 I - c v_a v_a^T: symmetric, the
               // upper triangle is accumulated and mirrored after the loop
                const PT dd = ga * (ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
                PT sa[3];
                for (int i = 0; i < 3; ++i) sa[i] = Aa[i] + Ba[i];
                for (int i = 0; i < 3; ++i)
                    for (int k = i; k < 3; ++k) {
                        K[0][i][k] = fma(sa[i], pa[k], K[0][i][k]);
                        if (L::HAS_N) K[0][i][k] = fma(-Va[i], va[k], K[0][i][k]);
                    }
                for (int i = 0; i < 3; ++i) K[0][i][i] += dd;
            }
#pragma unroll
            for (int j = 1; j < 5; ++j) {
                if (j < nb) {
                    const int b = (a + j) & 7;
                    PT pb[3], rb[3], vb[3] = {0, 0, 0};
                    for (int i = 0; i < 3; ++i) {
                        rb[i] = ldg(L::G, b, i, q);
                        pb[i] = L::TL ? ldg(L::P, b, i, q) : rb[i];
                    }
                    if (L::HAS_N) {
                        vb[0] = Nq[0] * pb[0] + Nq[3] * pb[1] + Nq[4] * pb[2];
                        vb[1] = Nq[3] * pb[0] + Nq[1] * pb[1] + Nq[5] * pb[2];
                        vb[2] = Nq[4] * pb[0] + Nq[5] * pb[1] + Nq[2] * pb[2];
                    }
                    const PT d = ga * (ra[0] * rb[0] + ra[1] * rb[1] + ra[2] * rb[2]);
                    for (int i = 0; i < 3; ++i)
                        for (int k = 0; k < 3; ++k) {
                            K[j][i][k] = fma(Aa[i], pb[k], K[j][i][k]);
                            K[j][i][k] = fma(Ba[k], pb[i], K[j][i][k]);
                            if (L::HAS_N) K[j][i][k] = fma(-Va[i], vb[k], K[j][i][k]);
                        }
                    for (int i = 0; i < 3; ++i) K[j][i][i] += d;
                }
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < i; ++k) K[0][i][k] = K[0][k][i];
    }
    // nodal forces straight from the lanes (element le's 24 values are contiguous over its 8 lanes)
    if (t == 0) s_el[le] = (int32_t)el;
    if (active)
        for (int i = 0; i < 3; ++i) R.fe[el * 24 + 3 * a + i] = fa[i];
    if (active && FORM == FORM_SS && MAT == M_J2)
        for (int s = 0; s < 7; ++s) A.state_t[((int64_t)s * NQ + t) * n + pos] = st_t[s];

    // records: staged through the (now free) quadrature-point area, 4 elements at a time, then written
    // contiguously per element (consecutive plan positions are mostly consecutive elements)
    VT *stage = reinterpret_cast<VT *>(s_qp);
    const int valid = (R.q_end - q0) < EPB ? (int)(R.q_end - q0) : EPB;
    for (int half = 0; half < 2; ++half) {
        lds_barrier();  // every lane is done with s_qp (phase 2) or with the previous half's copy-out
        if (active && (le >> 2) == half) {
            const int ls = le & 3;
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                if (j < nb) {
                    const int b = (a + j) & 7;
                    VT *blk = stage + ls * kRec8 + (a <= b ? block_off8(a, b) : block_off8(b, a));
                    if (j == 0) {  // diagonal block: upper triangle
                        for (int i = 0; i < 3; ++i)
                            for (int k = i; k < 3; ++k) blk[kSym6[i][k]] = (VT)K[0][i][k];
                    } else {
                        for (int i = 0; i < 3; ++i)
                            for (int k = 0; k < 3; ++k) {
                                if (a <= b)
                                    blk[3 * i + k] = (VT)K[j][i][k];
                                else
                                    blk[3 * k + i] = (VT)K[j][i][k];  // store K_ba = K_ab^T
                            }
                    }
                }
            }
        }
        lds_barrier();
        const int nv = valid - 4 * half;  // elements of this half
        // 16-byte copies: records start at multiples of 16 bytes and hold a multiple of 16 bytes
        constexpr int VW = 16 / (int)sizeof(VT);
        typedef VT nvec __attribute__((ext_vector_type(VW)));
        static_assert((kRec8 * (int)sizeof(VT)) % 16 == 0, "records must be whole 16-byte units");
        const int nw = (nv < 4 ? nv : 4) * (kRec8 / VW);
        for (int w = lane; w < nw; w += 64) {
            const int l = w / (kRec8 / VW);
            // non-temporal: the records are read back by the gather only after the whole pass, so caching
            // them would only evict the pass's inputs
            const nvec val = reinterpret_cast<const nvec *>(stage)[w];
            __builtin_nontemporal_store(val, reinterpret_cast<nvec *>(reinterpret_cast<VT *>(R.kbuf) +
                                                                 (int64_t)s_el[4 * half + l] * kRec8 +
                                                                 (w - l * (kRec8 / VW)) * VW));
        }
    }
}

// ------------------------------------------------------------------------------------------------
// C3D8 / C3D8TL, pass 2: two row nodes per 64-lane workgroup, one per 32-lane half. Sums the node's rows of
// every incident element's record in ascending element order (the order of the reference's serial
// VIJ -> CSR scatter, csrgenerator.pyx:112-113, and P/F scatter, mk2.pyx:187-193) and writes the node's 3
// CSR rows (contiguous: the rows of one node have equal length) and P, F once. A row node's inputs are one
// descriptor and one blob (incident elements + contribution lists), both read in gather order, so a node
// costs three dependent global loads: descriptor, blob, element blocks.
// ------------------------------------------------------------------------------------------------
// A row node's 3 CSR rows are n = 3 L contiguous values. With FP64 values the gathers stage them in LDS shifted by
// row_shift(row) values, so that a value's LDS index and its global index agree modulo the 16-byte vector width,
// and write them in 16-byte units with the unaligned ends value by value (non-temporal: the CSR values are not read
// again in the call); lane l of nl >= 8 lanes takes part. FP32 values are stored value by value: staging them was
// measured slower (the rows are half the bytes, the two extra barriers cost more than the merged lines gain).
template <typename VT>
__device__ __forceinline__ int row_shift(const VT *row)
{
    return (int)((reinterpret_cast<uintptr_t>(row) / sizeof(VT)) % (16 / sizeof(VT)));
}

template <typename VT>
__device__ __forceinline__ void write_rows(VT *row, const VT *st, int n, int l, int nl)
{
    constexpr int VW = 16 / (int)sizeof(VT);
    typedef VT nvec __attribute__((ext_vector_type(VW)));
    const int head = (VW - row_shift(row)) % VW;  // values before the first 16-byte boundary
    const int nvecs = (n - head) / VW;            // whole vectors after it
    const int tail0 = head + nvecs * VW;          // first value after the last whole vector
    for (int w = l; w < nvecs; w += nl)
        __builtin_nontemporal_store(*reinterpret_cast<const nvec *>(st + head + w * VW),
                                    reinterpret_cast<nvec *>(row + head + w * VW));
    if (l < head) __builtin_nontemporal_store(st[l], row + l);
    if (l >= VW && l - VW < n - tail0) __builtin_nontemporal_store(st[tail0 + l - VW], row + tail0 + l - VW);
}

constexpr int kMaxDeg8 = 64;
constexpr int kMaxInc8 = 31;
constexpr int kMaxBlobWords8 = kMaxInc8 + (kMaxDeg8 + 1 + 8 * kMaxInc8 + 3) / 4;
constexpr int kBlk = 10;  // doubles per staged 3x3 block (16-byte aligned: the accumulation reads use b128 loads)

template <typename VT>
__global__ __launch_bounds__(64) void k_gather8(const int4 *__restrict__ gdesc, const int32_t *__restrict__ gblob,
                                                const VT *__restrict__ kbuf, const double *__restrict__ fe,
                                                VT *__restrict__ vals, double *__restrict__ P, double *__restrict__ F,
                                                int64_t q_begin, int64_t count, int64_t xcd_chunk)
{
    __shared__ __align__(16) double s_blk[2][64 * kBlk];  // per half: blocks K_ab (oriented for node a) of 8 elements
    __shared__ double s_fe[2][8 * 3];
    __shared__ int32_t s_blob[2][kMaxBlobWords8 + 1];
    __shared__ int32_t s_nk[2];
    // workgroups are dealt to the 8 XCDs round robin: XCD x takes the contiguous run of node pairs
    // [x * xcd_chunk, (x + 1) * xcd_chunk) of this launch's range rorder[q_begin, q_begin + count), so row nodes
    // that are close in the gather order (and share elements) share that XCD's L2
    const int64_t k = blockIdx.x >> 3;
    const int64_t pair = (int64_t)(blockIdx.x & 7) * xcd_chunk + k;
    if (k >= xcd_chunk || 2 * pair >= count) return;  // uniform over the workgroup
    const int lane = threadIdx.x, h = lane >> 5, hl = lane & 31;
    const int64_t idx = 2 * pair + h;
    const bool live = idx < count;  // the second half may lie past the end of the range
    const int4 ds = live ? gdesc[q_begin + idx] : make_int4(0, 0, 0, 0);
    const int32_t d = ds.x, base = ds.y, boff = ds.z;
    const int deg = ds.w & 255, nk = ds.w >> 8;
    const int L = 3 * deg;
    {
        const int nwords = nk + (deg + 1 + 8 * nk + 3) / 4;
        for (int w = hl; w < nwords; w += 32) s_blob[h][w] = gblob[boff + w];
        if (hl == 0) s_nk[h] = nk;
    }
    __syncthreads();
    const int nk_max = s_nk[0] > s_nk[1] ? s_nk[0] : s_nk[1];
    const int32_t *codes = s_blob[h];
    const uint8_t *cl = reinterpret_cast<const uint8_t *>(s_blob[h] + nk);
    const uint8_t *ent = cl + deg + 1;
    // outputs: lane hl owns the neighbour blocks of rank hl and hl + 32 (a hexahedral node has at most 64
    // neighbours here; 27 on structured meshes)
    double acc[2][9];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[u][t] = 0.0;
    double f = 0.0, fa = 0.0;
    int cur[2] = {0, 0}, cend[2] = {0, 0};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int m = hl + 32 * u;
        if (live && m < deg) {
            cur[u] = cl[m];
            cend[u] = cl[m + 1];
        }
    }
    for (int c0 = 0; c0 < nk_max; c0 += 8) {
        // 1) lane hl fetches blocks hl and hl + 32 of this chunk (element slot = block >> 3, node b = block & 7):
        //    all 64 loads of a node in flight at once
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const int blk = hl + 32 * rr;
            const int sl = blk >> 3, b = blk & 7;
            if (live && c0 + sl < nk) {
                const int32_t code = codes[c0 + sl];
                const int64_t e = code >> 3;
                const int a = code & 7;
                double v[9];
                load_block8(kbuf + e * kRec8, a, b, v);
                double *dst = s_blk[h] + blk * kBlk;
#pragma unroll
                for (int t = 0; t < 9; ++t) dst[t] = v[t];
                if (b < 3) s_fe[h][sl * 3 + b] = fe[e * 24 + 3 * a + b];
            }
        }
        __syncthreads();
        // 2) the owner of neighbour block m adds its contributions of this chunk, in ascending element order
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            for (; cur[u] < cend[u]; ++cur[u]) {
                const int e8 = ent[cur[u]];
                const int sl = (e8 >> 3) - c0;
                if (sl >= 8) break;
                const double *src = s_blk[h] + (sl * 8 + (e8 & 7)) * kBlk;
#pragma unroll
                for (int t = 0; t < 9; ++t) acc[u][t] += src[t];
            }
        }
        if (live && hl < 3) {
            const int nc = (nk - c0) < 8 ? (nk - c0) : 8;
            for (int sl = 0; sl < nc; ++sl) {
                const double x = s_fe[h][sl * 3 + hl];
                f -= x;
                fa += fabs(x);
            }
        }
        __syncthreads();
    }
    // the node's 3 CSR rows: FP64 staged in its half's block area (free after the last chunk's barrier) and
    // written in 16-byte units
    VT *const row = vals + base;
    VT *const st = reinterpret_cast<VT *>(s_blk[h]) + row_shift(row);
    static_assert(16 / sizeof(VT) - 1 + 9 * kMaxDeg8 <= 64 * kBlk * sizeof(double) / sizeof(VT),
                  "a node's rows must fit its half's block area");
    constexpr bool staged = sizeof(VT) == 8;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int m = hl + 32 * u;
        if (live && m < deg) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    if (staged)
                        st[i * L + 3 * m + kk] = (VT)acc[u][3 * i + kk];
                    else
                        row[i * L + 3 * m + kk] = (VT)acc[u][3 * i + kk];
                }
        }
    }
    if constexpr (staged) {
        __syncthreads();
        if (live) write_rows(row, st, 3 * L, hl, 32);
    }
    if (live && hl < 3) {
        P[d + hl] = f;
        F[d + hl] = fa;
    }
}

// ------------------------------------------------------------------------------------------------
// C3D8 / C3D8TL, pass 2 on boxGen structured grids (boxgen.py:125-185: node (i0, i1, i2) has index
// (i0 N1 + i1) N2 + i2 and dofs 3 * index, element (e0, e1, e2) likewise, C3D8 local nodes at fixed corner
// offsets). Everything the descriptor gather reads from its lists follows from the indices here:
//   * neighbours of a node: offsets d in {-1, 0, 1}^3 inside the grid; the CSR columns are sorted, so the
//     rank of d is its lexicographic rank among the valid offsets;
//   * degree = c0 c1 c2 with c_k = number of valid offsets on axis k, and the CSR offset of the node's rows
//     = 9 x (sum of the degrees of all lower nodes), a product of per-axis prefix sums;
//   * the elements holding both the node and neighbour d are the incident elements (corner offset o) with
//     o_k = 0 where d_k = -1 and o_k = 1 where d_k = +1; ascending o is ascending element index.
// One row node per 32-lane half, lane m owns neighbour block m and loads its own contributions: a node costs
// one dependent load (its index in the gather order) before the element blocks. Sums run in ascending
// element order, as in the descriptor gather: both give bitwise the same CSR values, P and F.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int local_of_offset(int x, int y, int z)
{
    // inverse of the C3D8 corner table {(0,0,0),(0,0,1),(1,0,1),(1,0,0),(0,1,0),(0,1,1),(1,1,1),(1,1,0)}
    constexpr int lut = (0 << 0) | (1 << 3) | (4 << 6) | (5 << 9) | (3 << 12) | (2 << 15) | (7 << 18) | (6 << 21);
    return (lut >> (3 * (4 * x + 2 * y + z))) & 7;
}

// sum of c(j) = 1 + [j > 0] + [j < n - 1] over j < i, for a grid axis of n >= 2 nodes
__device__ __forceinline__ int64_t axis_prefix(int64_t i, int64_t n) { return i == 0 ? 0 : (i >= n ? 3 * n - 2 : 3 * i - 1); }

struct GridArgs {
    int64_t E0, E1, E2;  // elements per axis (slowest first)
};

template <typename VT>
__global__ __launch_bounds__(64) void k_gather8s(GridArgs G, const int32_t *__restrict__ rorder,
                                                 const VT *__restrict__ kbuf, const double *__restrict__ fe,
                                                 VT *__restrict__ vals, double *__restrict__ P,
                                                 double *__restrict__ F, double *__restrict__ d64, int64_t q_begin,
                                                 int64_t count, int64_t xcd_chunk)
{
    // per half: block (a, b) of the incident element at corner o in slot 8 o + b, in the records' precision (9
    // values of a slot of SB = 10 doubles or 12 floats, 16-byte multiples); the spare double at value FO of slot
    // 8 o + c holds component c of the element's nodal force f_a. LDS 10 KB (FP64: 16 workgroups per CU) or 6 KB
    // (FP32 records, exact in float: the sums below are FP64 either way)
    constexpr int SB = sizeof(VT) == 8 ? kBlk : 12;
    constexpr int FO = sizeof(VT) == 8 ? 9 : 10;
    __shared__ __align__(16) VT s_blk[2][64 * SB];
    auto fe_slot = [&](int hh, int slot) { return reinterpret_cast<double *>(s_blk[hh] + slot * SB + FO); };
    const int64_t k = blockIdx.x >> 3;
    const int64_t pair = (int64_t)(blockIdx.x & 7) * xcd_chunk + k;
    if (k >= xcd_chunk || 2 * pair >= count) return;  // uniform over the workgroup
    const int lane = threadIdx.x, h = lane >> 5, hl = lane & 31;
    const int64_t idx = 2 * pair + h;
    const bool live = idx < count;  // the second half may lie past the end of the range
    // node and element indices in 32 bits (the plan's CSR is int32-indexed, so every count here fits): 32-bit
    // division instead of the long 64-bit sequences
    const int32_t E0 = (int32_t)G.E0, E1 = (int32_t)G.E1, E2 = (int32_t)G.E2;
    const int32_t N1 = E1 + 1, N2 = E2 + 1, N0 = E0 + 1;
    const int32_t r = live ? rorder[q_begin + idx] : 0;
    const int32_t i0 = r / (N1 * N2), rem = r - i0 * (N1 * N2), i1 = rem / N2, i2 = rem - i1 * N2;
    const int lo0 = i0 > 0, lo1 = i1 > 0, lo2 = i2 > 0;
    const int32_t n12 = E1 * E2;
    auto elem = [&](int o, int64_t &e) {  // incident element at corner offset o (0..7); false outside the grid
        const int32_t e0 = i0 - 1 + (o >> 2), e1 = i1 - 1 + ((o >> 1) & 1), e2 = i2 - 1 + (o & 1);
        e = (int64_t)(e0 * n12 + e1 * E2 + e2);
        return e0 >= 0 && e0 < E0 && e1 >= 0 && e1 < E1 && e2 >= 0 && e2 < E2;
    };
    // 1) lane hl fetches blocks hl and hl + 32 (corner o = block >> 3, element node b = block & 7)
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int blk = hl + 32 * rr;
        const int o = blk >> 3, b = blk & 7;
        int64_t e;
        if (live && elem(o, e)) {
            const int a = local_of_offset(1 - (o >> 2), 1 - ((o >> 1) & 1), 1 - (o & 1));
            double v[9];
            load_block8(kbuf + e * kRec8, a, b, v);
            VT *dst = s_blk[h] + blk * SB;
#pragma unroll
            for (int t = 0; t < 9; ++t) dst[t] = (VT)v[t];  // v holds record values: exact
            if (b < 3) *fe_slot(h, blk) = fe[e * 24 + 3 * a + b];
        }
    }
    __syncthreads();
    const int c0 = 1 + lo0 + (i0 < N0 - 1), c1 = 1 + lo1 + (i1 < N1 - 1), c2 = 1 + lo2 + (i2 < N2 - 1);
    const int deg = c0 * c1 * c2, L = 3 * deg;
    const int m = hl;
    const int64_t base = 9 * (axis_prefix(i0, N0) * (3 * N1 - 2) * (3 * N2 - 2) +
                              (int64_t)c0 * axis_prefix(i1, N1) * (3 * N2 - 2) + (int64_t)c0 * c1 * axis_prefix(i2, N2));
    double acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = 0.0;
    if (live && m < deg) {
        // 2) neighbour offset of rank m; its contributions in ascending element (= corner) order
        const int d0 = m / (c1 * c2) - lo0, d1 = (m / c2) % c1 - lo1, d2 = m % c2 - lo2;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int o0 = o >> 2, o1 = (o >> 1) & 1, o2 = o & 1;
            int64_t e;
            if ((d0 == 0 || d0 == 2 * o0 - 1) && (d1 == 0 || d1 == 2 * o1 - 1) && (d2 == 0 || d2 == 2 * o2 - 1) &&
                elem(o, e)) {
                const int b = local_of_offset(1 - o0 + d0, 1 - o1 + d1, 1 - o2 + d2);
                const VT *src = s_blk[h] + (8 * o + b) * SB;
#pragma unroll
                for (int t = 0; t < 9; ++t) acc[t] += (double)src[t];
            }
        }
    }
    if (live && m >= 29) {  // P and F of component m - 29, in ascending element order
        const int c = m - 29;
        double f = 0.0, fa = 0.0;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            int64_t e;
            if (elem(o, e)) {
                const double x = *fe_slot(h, 8 * o + c);
                f -= x;
                fa += fabs(x);
            }
        }
        P[3 * r + c] = f;
        F[3 * r + c] = fa;
    }
    // EWFE_F32_D64: E = -(sum of the rounded off-diagonal blocks) - (rounded diagonal block). The rounded blocks go
    // through the half's block area (free once every lane has summed its contributions and the forces), then lane
    // t < 9 sums entry t over the neighbours in ascending order: sums of FP32 values, exact in FP64 unless their
    // magnitudes spread over more than ~24 binades, so the separate rebuild pass gives the same corrections
    if (d64) {  // uniform over the grid; only FP32 values carry corrections, so the rounded blocks are floats
        float *st = reinterpret_cast<float *>(s_blk[h]);  // 32 x 9 floats (lane stride 9 words: no bank conflicts)
        static_assert(32 * 9 * sizeof(float) <= 64 * SB * sizeof(VT), "the rounded blocks must fit the block area");
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 9; ++t) st[9 * m + t] = (live && m < deg) ? (float)acc[t] : 0.f;  // acc = 0 beyond
        __syncthreads();
        if (live && m < 9) {
            const int dr = (lo0 * c1 + lo1) * c2 + lo2;  // rank of the node itself (offset 0, 0, 0)
            double x = 0.0;
#pragma unroll
            for (int k = 0; k < 32; ++k) {  // every load issued at once: no branch, the diagonal masked by a select
                const double v = (double)st[9 * k + m];
                x += k == dr ? 0.0 : v;
            }
            d64[9 * r + m] = -x - (double)st[9 * dr + m];
        }
    }
    // 3) the node's 3 CSR rows: FP64 staged in its half's block area and written in 16-byte units
    VT *const row = vals + base;
    if constexpr (sizeof(VT) == 8) {
        static_assert(16 / sizeof(VT) - 1 + 3 * 3 * 27 <= 64 * SB, "a node's rows must fit its half's block area");
        VT *const st = s_blk[h] + row_shift(row);
        __syncthreads();  // every lane is done reading the blocks and forces
        if (live && m < deg)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) st[i * L + 3 * m + kk] = (VT)acc[3 * i + kk];
        __syncthreads();
        if (live) write_rows(row, st, 3 * L, hl, 32);
    } else if (live && m < deg) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) __builtin_nontemporal_store((VT)acc[3 * i + kk], row + i * L + 3 * m + kk);
    }
}

// ------------------------------------------------------------------------------------------------
// Body force (displacementelement/element.py:348-371, assembled by nonlinearimplicitstatic.py:903-945):
// P_e,a += sum_q N_a(xi_q) b detJ_q w_q for the elements of the load's element set. The load vector b is
// constant over the set, so each element contributes b times its nodal volume weights
// W_e,a = sum_q N_a(xi_q) detJ_q w_q (reference configuration); the second kernel sums them per node in
// ascending element order and adds b times the sum to PExt.
// ------------------------------------------------------------------------------------------------
template <int NN, int NQ>
__global__ __launch_bounds__(256) void k_body_weights(const int32_t *__restrict__ edof, const int32_t *__restrict__ order,
                                                      const double *__restrict__ X, const uint8_t *__restrict__ mask,
                                                      double *__restrict__ W, int64_t n_el)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_el) return;
    const int64_t e = order[p];
    double w[NN];
#pragma unroll
    for (int a = 0; a < NN; ++a) w[a] = 0.0;
    if (!mask || mask[e]) {
        double x[NN][3];
#pragma unroll
        for (int a = 0; a < NN; ++a) {
            const int32_t d = edof[(int64_t)a * n_el + p];
#pragma unroll
            for (int c = 0; c < 3; ++c) x[a][c] = X[d + c];
        }
        for (int q = 0; q < NQ; ++q) {
            const DNRow dN = qp_dN<NN, NQ>(q);
            const double *Nq = qp_N<NN, NQ>(q);
            const double wq = qp_w<NN, NQ>(q);
            double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int a = 0; a < NN; ++a)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * x[a][j];
            const double wd = det3(J) * wq;
#pragma unroll
            for (int a = 0; a < NN; ++a) w[a] += Nq[a] * wd;
        }
    }
#pragma unroll
    for (int a = 0; a < NN; ++a) W[e * NN + a] = w[a];
}

__global__ __launch_bounds__(256) void k_body_nodes(const int32_t *__restrict__ rownode_dof,
                                                    const int32_t *__restrict__ ninc_ptr,
                                                    const int32_t *__restrict__ ninc, const double *__restrict__ W,
                                                    int npe, double b0, double b1, double b2,
                                                    double *__restrict__ PExt, int64_t n_rownodes)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rownodes) return;
    double v = 0.0;
    for (int32_t k = ninc_ptr[r]; k < ninc_ptr[r + 1]; ++k) {
        const int32_t code = ninc[k];  // 32 * element + local node, ascending element
        v += W[(int64_t)(code >> 5) * npe + (code & 31)];
    }
    const int32_t d = rownode_dof[r];
    PExt[d + 0] += v * b0;
    PExt[d + 1] += v * b1;
    PExt[d + 2] += v * b2;
}

// ------------------------------------------------------------------------------------------------
// Distributed load of type pressure (nonlinearimplicitstatic.py:460-501; the element side is the Marmot
// element's computeDistributedLoad, element.pxd:97-144): for the listed elements, face `face` (Abaqus numbering:
// S1 zeta = -1, S2 zeta = +1, S3 eta = -1, S4 xi = +1, S5 eta = +1, S6 xi = -1) carries the traction -p n on
// the reference configuration: P_a = -p sum_g N_a n dA w_g over a 2 x 2 (8 nodes) or 3 x 3 (20 nodes) Gauss
// rule on the face. Per element [npe][3] into scratch, then summed per node in ascending element order.
// ------------------------------------------------------------------------------------------------
template <int NN>
__global__ __launch_bounds__(256) void k_face_pressure(const int32_t *__restrict__ edof,
                                                       const int32_t *__restrict__ pos_of_el,
                                                       const double *__restrict__ X, const int32_t *__restrict__ els,
                                                       int64_t n_list, int face, double pressure,
                                                       double *__restrict__ W, int32_t *__restrict__ slot_of_el,
                                                       int64_t n_el)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_list) return;
    const int64_t e = els[k];
    if (e < 0 || e >= n_el) return;  // not an element of the plan
    slot_of_el[e] = (int32_t)k;  // an element listed twice keeps one of its (equal) slots
    const int64_t p = pos_of_el[e];
    double x[NN][3];
    for (int a = 0; a < NN; ++a) {
        const int32_t d = edof[(int64_t)a * n_el + p];
        for (int c = 0; c < 3; ++c) x[a][c] = X[d + c];
    }
    // fixed parametric axis and its value; (u, v) = the cyclic successors of the fixed axis
    constexpr int kAxis[7] = {0, 2, 2, 1, 0, 1, 0};
    constexpr int kSign[7] = {0, -1, 1, -1, 1, 1, -1};
    const int ax = kAxis[face], sg = kSign[face];
    const int u = (ax + 1) % 3, v = (ax + 2) % 3;
    const int ng = NN == 8 ? 2 : 3;
    const double g2 = 0.5773502691896257645, g3 = 0.7745966692414833770;
    const double pts2[2] = {-g2, g2}, pts3[3] = {-g3, 0.0, g3}, wts3[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double f[NN][3];
    for (int a = 0; a < NN; ++a) f[a][0] = f[a][1] = f[a][2] = 0.0;
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j) {
            double xi[3];
            xi[ax] = sg;
            xi[u] = ng == 2 ? pts2[i] : pts3[i];
            xi[v] = ng == 2 ? pts2[j] : pts3[j];
            const double w = ng == 2 ? 1.0 : wts3[i] * wts3[j];
            double N[NN], dN[NN][3];
            shape8_20<NN>(xi, N, dN);
            double tu[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
            for (int a = 0; a < NN; ++a)
                for (int c = 0; c < 3; ++c) {
                    tu[c] += dN[a][u] * x[a][c];
                    tv[c] += dN[a][v] * x[a][c];
                }
            const double nA[3] = {sg * (tu[1] * tv[2] - tu[2] * tv[1]), sg * (tu[2] * tv[0] - tu[0] * tv[2]),
                                  sg * (tu[0] * tv[1] - tu[1] * tv[0])};
            for (int a = 0; a < NN; ++a)
                for (int c = 0; c < 3; ++c) f[a][c] -= pressure * N[a] * nA[c] * w;
        }
    for (int a = 0; a < NN; ++a)
        for (int c = 0; c < 3; ++c) W[(k * NN + a) * 3 + c] = f[a][c];
}

__global__ __launch_bounds__(256) void k_node_vec_sum(const int32_t *__restrict__ rownode_dof,
                                                      const int32_t *__restrict__ ninc_ptr,
                                                      const int32_t *__restrict__ ninc, const double *__restrict__ W,
                                                      const int32_t *__restrict__ slot_of_el, int npe,
                                                      double *__restrict__ PExt, int64_t n_rownodes)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rownodes) return;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int32_t k = ninc_ptr[r]; k < ninc_ptr[r + 1]; ++k) {
        const int32_t code = ninc[k];  // 32 * element + local node, ascending element
        const int32_t slot = slot_of_el[code >> 5];
        if (slot < 0) continue;  // element not loaded
        const double *w = W + ((int64_t)slot * npe + (code & 31)) * 3;
        v0 += w[0];
        v1 += w[1];
        v2 += w[2];
    }
    const int32_t d = rownode_dof[r];
    PExt[d + 0] += v0;
    PExt[d + 1] += v1;
    PExt[d + 2] += v2;
}

// ------------------------------------------------------------------------------------------------
// Quadrature-point results (the per-point "stress" and "strain" arrays the reference's elements expose through
// getResultArray, collected by ElementResultCollector, edelweissfe/utils/elementresultcollector.pyx:37-127),
// at a given displacement U and state, one thread per (element, point), element order out[e][q][6]:
//   small strain (displacementelement/element.py:222-235, 290-346): strain = eps(U), engineering Voigt
//     [11, 22, 33, 12, 13, 23]; stress = C : eps (linear elastic) or C : (eps - eps_p) with the plastic strain of
//     `state` (J2, the stress of a converged state);
//   total Lagrangian (displacementtlelement/element.py:377-412): strain = Green-Lagrange E, stress = Kirchhoff
//     tau (Neo-Hooke Wa), both Voigt [11, 22, 33, 12, 23, 13], engineering shear strains.
// ------------------------------------------------------------------------------------------------
template <int NN, int NQ, int FORM, int MAT>
__global__ __launch_bounds__(256) void k_qp_results(EvalArgs A, const double *__restrict__ state, int what,
                                                    double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = A.n_el;
    if (t >= n * NQ) return;
    const int64_t p = t / NQ;
    const int q = (int)(t % NQ);
    const int64_t e = A.order[p];
    const DNRow dN = qp_dN<NN, NQ>(q);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ji[3][3], H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    int32_t dof[NN];
    for (int a = 0; a < NN; ++a) dof[a] = A.edof[(int64_t)a * n + p];
    for (int a = 0; a < NN; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * A.X[dof[a] + j];
    inv3(J, Ji);
    for (int a = 0; a < NN; ++a) {
        double g[3];
        for (int j = 0; j < 3; ++j) g[j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[i][j] += A.U[dof[a] + i] * g[j];
    }
    double v[6];
    const MatParams M = mat_at(A, p);
    const double Epar = A.el_param ? A.el_param[e] : M.p[0];
    if (FORM == FORM_SS) {
        const double eps[6] = {H[0][0], H[1][1], H[2][2], H[0][1] + H[1][0], H[0][2] + H[2][0], H[1][2] + H[2][1]};
        if (what == 1) {
            for (int i = 0; i < 6; ++i) v[i] = eps[i];
        } else {
            double ee[6];
            for (int i = 0; i < 6; ++i) ee[i] = eps[i];
            if (MAT == M_J2)
                for (int s = 0; s < 6; ++s) ee[s] -= state[((int64_t)s * NQ + q) * n + p];
            const double nu = M.p[1];
            const double f = Epar / ((1 + nu) * (1 - 2 * nu));
            const double lam = f * nu, mu = f * (1 - 2 * nu) / 2;
            const double tr = ee[0] + ee[1] + ee[2];
            for (int i = 0; i < 3; ++i) v[i] = lam * tr + 2 * mu * ee[i];
            for (int i = 3; i < 6; ++i) v[i] = mu * ee[i];
        }
    } else {
        double F[3][3], Fi[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
        if (what == 1) {
            double E[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    E[i][j] = 0.5 * (H[i][j] + H[j][i] + H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
            const double o[6] = {E[0][0], E[1][1], E[2][2], 2 * E[0][1], 2 * E[1][2], 2 * E[0][2]};
            for (int i = 0; i < 6; ++i) v[i] = o[i];
        } else {
            const double Jd = inv3(F, Fi);
            Recipe r;
            if (MAT == M_SVK)
                svk_recipe(Epar, M.p[1], F, r);
            else
                nh_recipe(Epar, M.p[1], F, Jd, r);  // r.S: [11, 22, 33, 12, 13, 23]
            const double o[6] = {r.S[0], r.S[1], r.S[2], r.S[3], r.S[5], r.S[4]};
            for (int i = 0; i < 6; ++i) v[i] = o[i];
        }
    }
    for (int i = 0; i < 6; ++i) out[(e * NQ + q) * 6 + i] = v[i];
}

// ------------------------------------------------------------------------------------------------
// Adjoint: grad[e] = lambda^T dR/dtheta_e for a per-element material parameter theta_e (the first material
// property: mu of Neo-Hooke Wa on total-Lagrangian hexahedra, E of linear elasticity on small-strain
// hexahedra). R = P + PExt with P = -f_int, so grad[e] = -sum_a lambda_a . df_a/dtheta_e. The element
// kinematics are recomputed (nothing of the forward evaluation is stored).
//   Neo-Hooke Wa: tau = mu B + lamBar I, lamBar = (K - 2mu/3)(J^2 - J) - mu  =>
//                 dtau/dmu = B - ((2/3)(J^2 - J) + 1) I,  f_a = sum_q wd tau NA_a
//   linear elastic: sigma is linear in E  =>  dsigma/dE = sigma / E,  f_a = sum_q wd sigma g_a
// One 64-lane workgroup holds 8 elements, lane = (element, quadrature point); the 8 point contributions
// are summed with a fixed butterfly (deterministic).
// ------------------------------------------------------------------------------------------------
template <int FORM, int MAT>
__global__ __launch_bounds__(64) void k_adjoint8(EvalArgs A, const double *__restrict__ lam, double *__restrict__ grad)
{
    constexpr int NN = 8, EPB = 8;
    __shared__ double s_nd[9 * NN * EPB];  // [X, U, lambda][node][le]
    const int lane = threadIdx.x;
    const int le = lane >> 3, t = lane & 7;
    const int64_t pos = (int64_t)blockIdx.x * EPB + le;
    const bool active = pos < A.n_el;
    const int64_t n = A.n_el;
    if (active) {
        const int32_t dof = A.edof[(int64_t)t * n + pos];
        for (int c = 0; c < 3; ++c) {
            s_nd[((0 + c) * NN + t) * EPB + le] = A.X[dof + c];
            s_nd[((3 + c) * NN + t) * EPB + le] = A.U[dof + c];
            s_nd[((6 + c) * NN + t) * EPB + le] = lam[dof + c];
        }
    }
    __syncthreads();
    double contrib = 0.0;
    if (active) {
        const int q = t;
        auto nd = [&](int f, int a) { return s_nd[(f * NN + a) * EPB + le]; };
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ji[3][3];
        for (int a = 0; a < NN; ++a)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[i][j] += c_dN8[q][a][i] * nd(j, a);
        const double wd = c_w8[q] * inv3(J, Ji);
        double g[NN][3];
        double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < NN; ++a) {
            for (int j = 0; j < 3; ++j)
                g[a][j] = Ji[j][0] * c_dN8[q][a][0] + Ji[j][1] * c_dN8[q][a][1] + Ji[j][2] * c_dN8[q][a][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) H[i][j] += nd(3 + i, a) * g[a][j];
        }
        double D[3][3];  // d(stress)/d(theta), symmetric
        const MatParams M = mat_at(A, pos);
        double theta = M.p[0];
        if (A.el_param) theta = A.el_param[A.order[pos]];
        if (FORM == FORM_TL) {
            double F[3][3], Fi[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? 1.0 : 0.0);
            const double Jd = inv3(F, Fi);
            const double c = 2.0 / 3.0 * (Jd * Jd - Jd) + 1.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    D[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2] - (i == j ? c : 0.0);
            for (int a = 0; a < NN; ++a) {  // p_a = NA_a = F^-T g_a
                double p[3];
                for (int m = 0; m < 3; ++m) p[m] = g[a][0] * Fi[0][m] + g[a][1] * Fi[1][m] + g[a][2] * Fi[2][m];
                for (int i = 0; i < 3; ++i)
                    contrib += nd(6 + i, a) * wd * (D[i][0] * p[0] + D[i][1] * p[1] + D[i][2] * p[2]);
            }
        } else {
            Recipe r;
            le_recipe(M, theta, H, r);
            const double S[3][3] = {{r.S[0], r.S[3], r.S[4]}, {r.S[3], r.S[1], r.S[5]}, {r.S[4], r.S[5], r.S[2]}};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) D[i][j] = S[i][j] / theta;
            for (int a = 0; a < NN; ++a)
                for (int i = 0; i < 3; ++i)
                    contrib += nd(6 + i, a) * wd * (D[i][0] * g[a][0] + D[i][1] * g[a][1] + D[i][2] * g[a][2]);
        }
    }
    // sum over the 8 quadrature-point lanes of each element (fixed butterfly)
    contrib += __shfl_xor(contrib, 1, 8);
    contrib += __shfl_xor(contrib, 2, 8);
    contrib += __shfl_xor(contrib, 4, 8);
    if (active && t == 0) grad[A.order[pos]] = -contrib;
}

// ------------------------------------------------------------------------------------------------
// C3D20 / C3D20TL (27 points) and C3D20R / C3D20RTL (8 points): one element per 64-lane workgroup. The
// 27-point rule is evaluated in two chunks of points (14 + 13) so that the per-point data of a chunk fits
// 17 KB of LDS, which keeps 8 workgroups per CU (the register limit) instead of 5; the sums over the points
// keep their order.
// ------------------------------------------------------------------------------------------------
// REC: write the element's record (unique blocks + nodal forces) for the row-node gather instead of adding into
// the CSR values (then the launch needs no colouring)
// MF (FP32 records only): the tangent formed on the matrix cores (v_mfma_f32_16x16x4_f32) instead of FP64 FMAs
// (without MF, FP32 records hold the FP64 tangent rounded when it is stored).
template <int FORM, int MAT, typename VT, int NQ, bool REC, bool MF = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_hex20(EvalArgs A, RecordArgs R)
{
    static_assert(!MF || (REC && sizeof(VT) == 4 && MAT != M_SVK), "matrix-core tangent: FP32 records, rank form");
    constexpr int NN = 20;
    constexpr int QC = NQ > 16 ? 14 : NQ;  // points per chunk
    using L = QPLayout<NN, FORM>;
    __shared__ double s_nd[9 * NN];      // [field][node]
    // [field][point of the chunk]; with REC afterwards the staged record (written out in whole 16-byte units);
    // with MF the two staged halves of the record (T - C + S and Tb) and S_ab
    constexpr int kQP = L::NF * QC,
                  kStage = MF ? (2 * kRec20 + NN * NN) / 2 : (REC ? (kRec20 * (int)sizeof(VT) + 7) / 8 : 0);
    __shared__ __align__(16) double s_qp[kQP > kStage ? kQP : kStage];
    __shared__ int32_t s_dof[NN];
    __shared__ int32_t s_rs[3 * NN];     // CSR row starts of the element's 60 dof rows
    const int lane = threadIdx.x;
    const int64_t pos = A.p_begin + blockIdx.x;
    if (pos >= A.p_end) return;  // uniform over the workgroup
    const int64_t n = A.n_el;
    if (lane < NN) {
        const int32_t d = A.edof[(int64_t)lane * n + pos];
        s_dof[lane] = d;
        for (int c = 0; c < 3; ++c) {
            s_rs[3 * lane + c] = A.indptr[d + c];
            s_nd[(0 + c) * NN + lane] = A.X[d + c];
            s_nd[(3 + c) * NN + lane] = A.U[d + c];
            s_nd[(6 + c) * NN + lane] = A.dU[d + c];
        }
    }
    const MatParams M = mat_at(A, pos);
    double Epar = M.p[0];
    if (A.el_param) Epar = A.el_param[A.order[pos]];
    // lane = (a, g): the force row (a, i = g) and all three rows of the blocks K_ab, b = a + j (mod 20), for the
    // offsets j of group g: {0..3}, {4..7}, {8..10}, with j < nb (j = 10 only from a < 10: every unordered node pair
    // once). The lanes of one node share its point data; per point a lane reads node a's and its 4 partner nodes'
    // vectors from LDS once for 36 block entries.
    const bool row_lane = lane < 3 * NN;
    const int a = row_lane ? lane / 3 : 0, g = lane % 3, i = g;
    constexpr int kJ = NN / 2 + 1;  // 11 offsets
    constexpr int kG = 4;           // offsets per lane
    const int j0 = kG * g;
    const int nb = (a < NN / 2) ? kJ : kJ - 1;
    const int nj = nb - j0 < kG ? nb - j0 : kG;  // 4, 4, 3 or 2
    double K[kG][9];
    for (int j = 0; j < kG; ++j)
        for (int t = 0; t < 9; ++t) K[j][t] = 0.0;
    double fi = 0.0;
    // MF: the 60 x 60 element matrix padded to 64 x 64, its 10 upper 16 x 16 tiles (I <= J) accumulated over the
    // points as T - C (alpha p p^T - c v v^T) and Tb (beta p p^T), and S_ab = sum_q gamma (r_a . r_b) over the
    // 3 upper tiles of 32 x 32; lane (kq, c16) supplies row / column 16 I + c16 at inner index kq of each MFMA
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int kTiles = MF ? 10 : 1;
    f4 accT[kTiles], accB[kTiles], accS[MF ? 3 : 1];
    const int kq = lane >> 4, c16 = lane & 15;
    if constexpr (MF) {
        for (int t = 0; t < kTiles; ++t) accT[t] = accB[t] = (f4){0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < 3; ++t) accS[t] = (f4){0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll 1
    for (int q0 = 0; q0 < NQ; q0 += QC) {
        const int nq = NQ - q0 < QC ? NQ - q0 : QC;
        lds_barrier();  // the nodes are staged, or the previous chunk's points are summed
        // four lanes per point: lane (ql, s) = 4 ql + s; the groups of points past nq stay idle as a whole
        const int ql = lane >> 2, sq = lane & 3;
        if (ql < nq) {
            const int q = q0 + ql;
            auto nd = [&](int f, int b) { return s_nd[f * NN + b]; };
            auto put_g = [&](int b, int j, double v) { s_qp[(L::G + 3 * b + j) * QC + ql] = v; };
            auto put_p = [&](int b, int j, double v) { s_qp[(L::P + 3 * b + j) * QC + ql] = v; };
            Recipe r;
            double wd;
            const bool ok = qp_phase_coop<NN, FORM, MAT>(qp_dN<NN, NQ>(q), qp_w<NN, NQ>(q), sq, nd, put_g, put_p, A,
                                                         M, Epar, pos, q, NQ, r, wd);
            if (sq == 0) {  // every lane of the group holds the point's recipe; the first stores it
                if (!ok) atomic_min_pos(A.pnewdt, 0.5);
                auto put = [&](int f, double v) { s_qp[f * QC + ql] = v; };
                put(L::COEF + 0, wd * r.alpha);
                put(L::COEF + 1, wd * r.beta);
                put(L::COEF + 2, wd * r.gam);
                put(L::COEF + 3, wd * r.c);
                for (int m = 0; m < 6; ++m) put(L::NV + m, (MAT == M_J2 || MAT == M_SVK) ? r.N[m] : 0.0);
                for (int m = 0; m < 6; ++m) put(L::S + m, wd * r.S[m]);
            }
        }
        lds_barrier();
        auto ld = [&](int f, int q) { return s_qp[f * QC + q]; };
        if constexpr (MF) {
            // forces in FP64 (the FP64 variant's arithmetic): row lane (a, i)
            if (row_lane) {
                for (int q = 0; q < nq; ++q) {
                    const double S0 = ld(L::S + 0, q), S1 = ld(L::S + 1, q), S2 = ld(L::S + 2, q);
                    const double S3 = ld(L::S + 3, q), S4 = ld(L::S + 4, q), S5 = ld(L::S + 5, q);
                    const double Sr0 = (i == 0) ? S0 : (i == 1 ? S3 : S4);
                    const double Sr1 = (i == 0) ? S3 : (i == 1 ? S1 : S5);
                    const double Sr2 = (i == 0) ? S4 : (i == 1 ? S5 : S2);
                    fi += Sr0 * ld(L::P + 3 * a + 0, q) + Sr1 * ld(L::P + 3 * a + 1, q) + Sr2 * ld(L::P + 3 * a + 2, q);
                }
            }
            // T, Tb (and C): inner index = point kq of each group of 4 points of the chunk
#pragma unroll 1
            for (int k0 = 0; k0 < nq; k0 += 4) {
                const int q = k0 + kq;
                const bool qv = q < nq;
                const float al = qv ? (float)ld(L::COEF + 0, q) : 0.f;
                const float be = qv ? (float)ld(L::COEF + 1, q) : 0.f;
                float x[4], y[4] = {0.f, 0.f, 0.f, 0.f};
                float mc = 0.f;
                float Nv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (MAT == M_J2) {
                    mc = qv ? -(float)ld(L::COEF + 3, q) : 0.f;
                    for (int m = 0; m < 6; ++m) Nv[m] = qv ? (float)ld(L::NV + m, q) : 0.f;
                }
#pragma unroll
                for (int I = 0; I < 4; ++I) {
                    const int m = 16 * I + c16;
                    const bool mv = qv && m < 3 * NN;
                    x[I] = mv ? (float)ld(L::P + m, q) : 0.f;
                    if (MAT == M_J2 && mv) {  // v = N p, row m = (a, i) of the point's v
                        const int am = m / 3, im = m % 3;
                        const int j0 = im == 0 ? 0 : (im == 1 ? 3 : 4), j1 = im == 0 ? 3 : (im == 1 ? 1 : 5),
                                  j2 = im == 0 ? 4 : (im == 1 ? 5 : 2);
                        y[I] = Nv[j0] * (float)ld(L::P + 3 * am + 0, q) + Nv[j1] * (float)ld(L::P + 3 * am + 1, q) +
                               Nv[j2] * (float)ld(L::P + 3 * am + 2, q);
                    }
                }
                int t = 0;
#pragma unroll
                for (int I = 0; I < 4; ++I)
#pragma unroll
                    for (int J = I; J < 4; ++J, ++t) {
                        accT[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(al * x[I], x[J], accT[t], 0, 0, 0);
                        accB[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(be * x[I], x[J], accB[t], 0, 0, 0);
                        if (MAT == M_J2)
                            accT[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(mc * y[I], y[J], accT[t], 0, 0, 0);
                    }
            }
            // S: inner index = (point, component) kappa = 3 q + j, rows a = c16 (tile row 0) and 16 + c16 (row 1)
#pragma unroll 1
            for (int k0 = 0; k0 < 3 * nq; k0 += 4) {
                const int kap = k0 + kq, q = kap / 3, j = kap % 3;
                const bool qv = q < nq;
                const float gq = qv ? (float)ld(L::COEF + 2, q) : 0.f;
                const float r0 = qv ? (float)ld(L::G + 3 * c16 + j, q) : 0.f;
                const float r1 = (qv && 16 + c16 < NN) ? (float)ld(L::G + 3 * (16 + c16) + j, q) : 0.f;
                accS[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(gq * r0, r0, accS[0], 0, 0, 0);
                accS[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(gq * r0, r1, accS[1], 0, 0, 0);
                accS[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(gq * r1, r1, accS[2], 0, 0, 0);
            }
            continue;
        }
        if (!row_lane) continue;
#pragma unroll 1
        for (int q = 0; q < nq; ++q) {
            const double al = ld(L::COEF + 0, q), be = ld(L::COEF + 1, q), ga = ld(L::COEF + 2, q),
                         cc = ld(L::COEF + 3, q);
            double pa[3], ra[3], Nq[6], va[3] = {0.0, 0.0, 0.0};
            for (int m = 0; m < 3; ++m) {
                ra[m] = ld(L::G + 3 * a + m, q);
                pa[m] = ld(L::P + 3 * a + m, q);
            }
            {
                double Srow[3];
                const double S0 = ld(L::S + 0, q), S1 = ld(L::S + 1, q), S2 = ld(L::S + 2, q);
                const double S3 = ld(L::S + 3, q), S4 = ld(L::S + 4, q), S5 = ld(L::S + 5, q);
                Srow[0] = (i == 0) ? S0 : (i == 1 ? S3 : S4);
                Srow[1] = (i == 0) ? S3 : (i == 1 ? S1 : S5);
                Srow[2] = (i == 0) ? S4 : (i == 1 ? S5 : S2);
                fi += Srow[0] * pa[0] + Srow[1] * pa[1] + Srow[2] * pa[2];
            }
            if constexpr (MAT == M_SVK) {
                // St. Venant-Kirchhoff: K_ab,ik = alpha q_a,i q_b,k + beta q_a,k q_b,i + beta (r_a . r_b) B_ik
                // + delta_ik (p_a . S p_b), q = B p (alpha = lambda wd, beta = mu wd, S = wd tau)
                double Bm[3][3], Sm[3][3];
                {
                    const double b0 = ld(L::NV + 0, q), b1 = ld(L::NV + 1, q), b2 = ld(L::NV + 2, q);
                    const double b3 = ld(L::NV + 3, q), b4 = ld(L::NV + 4, q), b5 = ld(L::NV + 5, q);
                    Bm[0][0] = b0, Bm[1][1] = b1, Bm[2][2] = b2;
                    Bm[0][1] = Bm[1][0] = b3, Bm[0][2] = Bm[2][0] = b4, Bm[1][2] = Bm[2][1] = b5;
                    const double S0 = ld(L::S + 0, q), S1 = ld(L::S + 1, q), S2 = ld(L::S + 2, q);
                    const double S3 = ld(L::S + 3, q), S4 = ld(L::S + 4, q), S5 = ld(L::S + 5, q);
                    Sm[0][0] = S0, Sm[1][1] = S1, Sm[2][2] = S2;
                    Sm[0][1] = Sm[1][0] = S3, Sm[0][2] = Sm[2][0] = S4, Sm[1][2] = Sm[2][1] = S5;
                }
                double qa[3], tpa[3];
                for (int m = 0; m < 3; ++m) {
                    qa[m] = Bm[m][0] * pa[0] + Bm[m][1] * pa[1] + Bm[m][2] * pa[2];
                    tpa[m] = Sm[m][0] * pa[0] + Sm[m][1] * pa[1] + Sm[m][2] * pa[2];
                }
#pragma unroll
                for (int jj = 0; jj < kG; ++jj) {
                    if (jj < nj) {
                        const int b = (a + j0 + jj) % NN;
                        double pb[3], rb[3], qb[3];
                        for (int m = 0; m < 3; ++m) {
                            rb[m] = ld(L::G + 3 * b + m, q);
                            pb[m] = ld(L::P + 3 * b + m, q);
                        }
                        for (int m = 0; m < 3; ++m) qb[m] = Bm[m][0] * pb[0] + Bm[m][1] * pb[1] + Bm[m][2] * pb[2];
                        const double rr = be * (ra[0] * rb[0] + ra[1] * rb[1] + ra[2] * rb[2]);
                        const double d1 = tpa[0] * pb[0] + tpa[1] * pb[1] + tpa[2] * pb[2];
#pragma unroll
                        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                double v = al * qa[ii] * qb[k] + be * qa[k] * qb[ii] + rr * Bm[ii][k];
                                if (ii == k) v += d1;
                                K[jj][3 * ii + k] += v;
                            }
                    }
                }
                continue;
            }
            if (MAT == M_J2) {
                for (int m = 0; m < 6; ++m) Nq[m] = ld(L::NV + m, q);
                va[0] = cc * (Nq[0] * pa[0] + Nq[3] * pa[1] + Nq[4] * pa[2]);
                va[1] = cc * (Nq[3] * pa[0] + Nq[1] * pa[1] + Nq[5] * pa[2]);
                va[2] = cc * (Nq[4] * pa[0] + Nq[5] * pa[1] + Nq[2] * pa[2]);
            }
            double Aa[3], Ba[3];
            for (int m = 0; m < 3; ++m) {
                Aa[m] = al * pa[m];
                Ba[m] = be * pa[m];
            }
#pragma unroll
            for (int jj = 0; jj < kG; ++jj) {
                if (jj < nj) {
                    const int b = (a + j0 + jj) % NN;
                    double pb[3], rb[3];
                    for (int m = 0; m < 3; ++m) {
                        rb[m] = ld(L::G + 3 * b + m, q);
                        pb[m] = ld(L::P + 3 * b + m, q);
                    }
                    const double d = ga * (ra[0] * rb[0] + ra[1] * rb[1] + ra[2] * rb[2]);
                    double vb[3] = {0, 0, 0};
                    if (MAT == M_J2) {
                        vb[0] = Nq[0] * pb[0] + Nq[3] * pb[1] + Nq[4] * pb[2];
                        vb[1] = Nq[3] * pb[0] + Nq[1] * pb[1] + Nq[5] * pb[2];
                        vb[2] = Nq[4] * pb[0] + Nq[5] * pb[1] + Nq[2] * pb[2];
                    }
#pragma unroll
                    for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            double v = Aa[ii] * pb[k] + Ba[k] * pb[ii];
                            if (MAT == M_J2) v -= va[ii] * vb[k];
                            if (ii == k) v += d;
                            K[jj][3 * ii + k] += v;
                        }
                }
            }
        }
    }
    if constexpr (MF) {
        // record from the tiles: lane (kq, c16) holds rows 16 I + 4 kq + v (v = 0..3), column 16 J + c16 of tile
        // (I, J). Values (m, n), m <= n < 60, are the record's: block (a, b) = (m / 3, n / 3), a <= b. T - C (+ S_ab on
        // i == k) goes to the slot of K_ab[i][k] (i = m % 3, k = n % 3); Tb value (m, n) belongs to K_ab[k][i]. Each
        // slot is written by one lane in each half, and the halves are summed when the record is written out.
        lds_barrier();  // every lane is done with the point data
        float *sT = reinterpret_cast<float *>(s_qp), *sB = sT + kRec20, *sS = sB + kRec20;
        const int64_t el = A.order[pos];
        if (row_lane) R.fe[el * 60 + lane] = fi;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int rb = (u == 2 ? 16 : 0), cb = (u == 0 ? 0 : 16);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ra = rb + 4 * kq + v, cn = cb + c16;
                if (ra <= cn && cn < NN) sS[ra * NN + cn] = accS[u][v];
            }
        }
        lds_barrier();
        int t = 0;
#pragma unroll
        for (int I = 0; I < 4; ++I)
#pragma unroll
            for (int J = I; J < 4; ++J, ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = 16 * I + 4 * kq + v, n2 = 16 * J + c16;
                    if (m <= n2 && n2 < 3 * NN) {
                        const int am = m / 3, im = m % 3, bn = n2 / 3, kn = n2 % 3;
                        float val = accT[t][v];
                        if (im == kn) val += sS[am * NN + bn];
                        const bool offd = am < bn;
                        sT[offd ? block_off20(am, bn) + 3 * im + kn : block_off20(am, am) + kSym6[im][kn]] = val;
                        sB[offd ? block_off20(am, bn) + 3 * kn + im : block_off20(am, am) + kSym6[im][kn]] = accB[t][v];
                    }
                }
        lds_barrier();
        constexpr int NW = kRec20 / 4;
        static_assert(kRec20 % 4 == 0, "records are whole 16-byte units");
        f4 *dst = reinterpret_cast<f4 *>(reinterpret_cast<float *>(R.kbuf) + el * kRec20);
        for (int w = lane; w < NW; w += 64) {
            const f4 sum = reinterpret_cast<const f4 *>(sT)[w] + reinterpret_cast<const f4 *>(sB)[w];
            __builtin_nontemporal_store(sum, dst + w);
        }
        return;
    }
    if (REC) {
        // record of element el staged in LDS (the point data are no longer read once every lane passed the
        // barrier): the blocks as the record holds them (K_ab row-major for a < b, K_ba = K_ab^T otherwise, the
        // upper triangle of K_aa), then written out in whole 16-byte units; the nodal forces f_(a,i), lane 3a + i
        lds_barrier();
        VT *s_rec = reinterpret_cast<VT *>(s_qp);
        const int64_t el = A.order[pos];
        if (row_lane) {
            R.fe[el * 60 + lane] = fi;
#pragma unroll
            for (int jj = 0; jj < kG; ++jj) {
                if (jj < nj) {
                    const int j = j0 + jj, b = (a + j) % NN;
                    if (j == 0) {
#pragma unroll
                        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                            for (int k = ii; k < 3; ++k) s_rec[block_off20(a, a) + kSym6[ii][k]] = (VT)K[0][3 * ii + k];
                    } else if (a < b) {
#pragma unroll
                        for (int t = 0; t < 9; ++t) s_rec[block_off20(a, b) + t] = (VT)K[jj][t];
                    } else {
#pragma unroll
                        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                            for (int k = 0; k < 3; ++k) s_rec[block_off20(b, a) + 3 * k + ii] = (VT)K[jj][3 * ii + k];
                    }
                }
            }
        }
        lds_barrier();
        static_assert((kRec20 * (int)sizeof(VT)) % 16 == 0, "records are whole 16-byte units");
        constexpr int NW = kRec20 * (int)sizeof(VT) / 16;
        typedef float u4 __attribute__((ext_vector_type(4)));  // 16 bytes moved as they are
        const u4 *src = reinterpret_cast<const u4 *>(s_rec);
        u4 *dst = reinterpret_cast<u4 *>(reinterpret_cast<VT *>(R.kbuf) + el * kRec20);
        for (int w = lane; w < NW; w += 64) __builtin_nontemporal_store(src[w], dst + w);
        return;
    }
    if (!row_lane) return;
    const int32_t row = s_dof[a] + i;
    A.P[row] -= fi;
    A.F[row] += fabs(fi);
    VT *vals = reinterpret_cast<VT *>(A.vals);
#pragma unroll
    for (int jj = 0; jj < kG; ++jj) {
        if (jj < nj) {
            const int j = j0 + jj, b = (a + j) % NN;
            const int32_t cab = A.colpos[pos * (NN * NN) + a * NN + b];
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int k = 0; k < 3; ++k) add_to(vals, (int64_t)s_rs[3 * a + ii] + cab + k, K[jj][3 * ii + k]);
            if (j > 0) {  // K_ba[k][ii] = K_ab[ii][k] into row k of node b, column ii of node a
                const int32_t cba = A.colpos[pos * (NN * NN) + b * NN + a];
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        add_to(vals, (int64_t)s_rs[3 * b + k] + cba + ii, K[jj][3 * ii + k]);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// C3D20 family, pass 2: one row node per 64-lane workgroup. The node's incident elements (at most 8, ascending
// element order) and per neighbour rank m the (slot, element node) pairs contributing to block m come from the
// descriptor and blob built with the plan. Lane t first fetches blocks t, t + 64, t + 128 of the 20 nk blocks
// (slot, b) into LDS (oriented for the row node), then lane m sums neighbour block m (and m + 64) over its
// contributions in ascending element order and writes the 3 x 3 values into the node's 3 CSR rows; lanes 0..2 sum
// P and F. Every CSR value, P and F entry is written once: no colouring, no read-modify-write.
// ------------------------------------------------------------------------------------------------
constexpr int kMaxInc20 = 8;
constexpr int kMaxDeg20 = 128;
constexpr int kMaxBlobWords20 = kMaxInc20 + (kMaxDeg20 + 1 + 20 * kMaxInc20 + 3) / 4;

template <typename VT>
__global__ __launch_bounds__(64) void k_gather20(const int4 *__restrict__ gdesc, const int32_t *__restrict__ gblob,
                                                 const VT *__restrict__ kbuf, const double *__restrict__ fe,
                                                 VT *__restrict__ vals, double *__restrict__ P,
                                                 double *__restrict__ F, int64_t r_begin, int64_t count,
                                                 int64_t xcd_chunk)
{
    __shared__ __align__(16) double s_blk[20 * kMaxInc20 * kBlk];
    __shared__ double s_fe[kMaxInc20 * 3];
    __shared__ int32_t s_blob[kMaxBlobWords20 + 1];
    // workgroups are dealt round-robin to the 8 XCDs: XCD x takes the row nodes [x * xcd_chunk, (x + 1) *
    // xcd_chunk) of the range, so neighbouring row nodes (which read the same element records) share one L2
    const int64_t idx = blockIdx.x >> 3, xr = (int64_t)(blockIdx.x & 7) * xcd_chunk + idx;
    if (idx >= xcd_chunk || xr >= count) return;  // uniform over the workgroup
    const int64_t r = r_begin + xr;
    const int lane = threadIdx.x;
    const int4 ds = gdesc[r];
    const int32_t d = ds.x, base = ds.y, boff = ds.z;
    const int deg = ds.w & 255, nk = ds.w >> 8;
    const int nwords = nk + (deg + 1 + 20 * nk + 3) / 4;
    for (int w = lane; w < nwords; w += 64) s_blob[w] = gblob[boff + w];
    __syncthreads();
    const int32_t *codes = s_blob;
    const uint8_t *cs = reinterpret_cast<const uint8_t *>(s_blob + nk);
    const uint8_t *ent = cs + deg + 1;
    for (int t = lane; t < 20 * nk; t += 64) {
        const int sl = t / 20, b = t % 20;
        const int64_t e = codes[sl] >> 5;
        const int a = codes[sl] & 31;
        double v[9];
        load_block20(kbuf + e * kRec20, a, b, v);
        double *dst = s_blk + t * kBlk;
#pragma unroll
        for (int u = 0; u < 9; ++u) dst[u] = v[u];
        if (b < 3) s_fe[sl * 3 + b] = fe[e * 60 + 3 * a + b];
    }
    __syncthreads();
    const int L = 3 * deg;
    double acc[2][9];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[u][t] = 0.0;
        const int m = lane + 64 * u;
        if (m < deg) {
            for (int c = cs[m]; c < cs[m + 1]; ++c) {
                const int code = ent[c];
                const double *src = s_blk + ((code >> 5) * 20 + (code & 31)) * kBlk;
#pragma unroll
                for (int t = 0; t < 9; ++t) acc[u][t] += src[t];
            }
        }
    }
    // the node's 3 CSR rows: FP64 staged in the block area and written in 16-byte units
    VT *const row = vals + base;
    constexpr bool staged = sizeof(VT) == 8;
    VT *const st = reinterpret_cast<VT *>(s_blk) + row_shift(row);
    static_assert(16 / sizeof(VT) - 1 + 9 * kMaxDeg20 <= 20 * kMaxInc20 * kBlk * sizeof(double) / sizeof(VT),
                  "a node's rows must fit the block area");
    if constexpr (staged) __syncthreads();  // every lane is done reading the blocks
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int m = lane + 64 * u;
        if (m < deg)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (staged)
                        st[i * L + 3 * m + k] = (VT)acc[u][3 * i + k];
                    else
                        __builtin_nontemporal_store((VT)acc[u][3 * i + k], row + i * L + 3 * m + k);
                }
    }
    if constexpr (staged) {
        __syncthreads();
        write_rows(row, st, 3 * L, lane, 64);
    }
    if (lane < 3) {  // P and F of component lane, in ascending element order
        double f = 0.0, fa = 0.0;
        for (int sl = 0; sl < nk; ++sl) {
            const double x = s_fe[sl * 3 + lane];
            f -= x;
            fa += fabs(x);
        }
        P[d + lane] = f;
        F[d + lane] = fa;
    }
}

// ------------------------------------------------------------------------------------------------
// small helper kernels
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *p, int64_t n, T v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename VT>
__global__ void k_dirichlet_rows(const int32_t *rows, int64_t n_rows, const int32_t *indptr, const int32_t *indices,
                                 VT *vals)
{
    // one wavefront per Dirichlet row
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    if (r >= n_rows) return;
    const int32_t row = rows[r];
    for (int32_t k = indptr[row] + (threadIdx.x & 63); k < indptr[row + 1]; k += 64)
        vals[k] = (indices[k] == row) ? (VT)1 : (VT)0;
}

template <typename T>
static int fill(T *p, int64_t n, T v, hipStream_t s)
{
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    HIP_CHECK(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
static int npe_of(int k) { return (k == EWFE_C3D8 || k == EWFE_C3D8TL) ? 8 : 20; }
static int nqp_of(int k) { return (k == EWFE_C3D20 || k == EWFE_C3D20TL) ? 27 : 8; }
static bool is_tl(int k) { return k == EWFE_C3D8TL || k == EWFE_C3D20TL || k == EWFE_C3D20RTL; }

extern "C" int ewfe_state_size(int mat_kind)
{
    switch (mat_kind) {
    case EWFE_LINEAR_ELASTIC: return 0;
    case EWFE_VON_MISES: return 7;
    case EWFE_NEO_HOOKE_WA: return 0;
    case EWFE_ST_VENANT_KIRCHHOFF: return 0;
    default: ewfe::set_error("unknown material kind"); return EWFE_EINVAL;
    }
}

static int alloc_records(ewfe_plan *p);
static int ensure_pipeline(ewfe_plan *p);

static int plan_create(ewfe_plan **out, int elem_kind, int64_t n_el, const int32_t *conn_host, int64_t n_nodes,
                       const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out, bool upload, int chunks)
{
    if (!out || !conn_host || !node_dof_host || n_el <= 0 || n_nodes <= 0 || n_dof <= 0) {
        ewfe::set_error("ewfe_plan_create: invalid arguments");
        return EWFE_EINVAL;
    }
    if (elem_kind < EWFE_C3D8 || elem_kind > EWFE_C3D20RTL) {
        ewfe::set_error("ewfe_plan_create: unknown element kind");
        return EWFE_EINVAL;
    }
    *out = nullptr;
    ewfe_plan *p = new (std::nothrow) ewfe_plan();
    if (!p) {
        ewfe::set_error("out of host memory");
        return EWFE_ENOMEM;
    }
    p->elem_kind = elem_kind;
    p->npe = npe_of(elem_kind);
    p->nqp = nqp_of(elem_kind);
    p->n_el = n_el;
    p->n_nodes = n_nodes;
    p->n_dof = n_dof;
    p->n_chunks = chunks > 0 ? chunks : 0;
    if ((int64_t)p->npe * p->npe * n_el >= ((int64_t)1 << 40)) {
        delete p;
        ewfe::set_error("mesh too large");
        return EWFE_ERANGE;
    }
    std::vector<int32_t> edof, colpos;
    ewfe::GatherHost gh;
    int rc;
    try {
        rc = ewfe::build_plan_host(*p, conn_host, node_dof_host, edof, colpos, gh);
    } catch (const std::bad_alloc &) {
        rc = EWFE_ENOMEM;
        ewfe::set_error("out of host memory while building the plan");
    }
    if (rc) {
        delete p;
        return rc;
    }
    if (!upload) {
        if (nnz_out) *nnz_out = p->nnz;
        *out = p;
        return 0;
    }
    auto fail = [&](hipError_t e, const char *what) {
        ewfe::set_error(std::string(what) + ": " + hipGetErrorString(e));
        ewfe_plan_destroy(p);
        return EWFE_EHIP;
    };
    hipError_t e;
    if ((e = hipGetDevice(&p->device)) != hipSuccess) return fail(e, "hipGetDevice");
    if ((rc = upload_tables(p->device))) {
        ewfe_plan_destroy(p);
        return rc;
    }
    std::vector<int32_t> order32(n_el);
    for (int64_t i = 0; i < n_el; ++i) order32[i] = (int32_t)p->order[i];
    struct Up {
        int32_t **dst;
        const void *src;
        size_t bytes;
    } ups[] = {
        {&p->d_indptr, p->indptr.data(), sizeof(int32_t) * (n_dof + 1)},
        {&p->d_indices, p->indices.data(), sizeof(int32_t) * p->nnz},
        {&p->d_edof, edof.data(), sizeof(int32_t) * edof.size()},
        {&p->d_colpos, colpos.data(), sizeof(int32_t) * colpos.size()},
        {&p->d_order, order32.data(), sizeof(int32_t) * n_el},
    };
    for (auto &u : ups) {
        if ((e = hipMalloc((void **)u.dst, u.bytes ? u.bytes : 4)) != hipSuccess) return fail(e, "hipMalloc");
        if (u.bytes && (e = hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy");
    }
    {
        Up rups[] = {
            {&p->d_rownode_dof, gh.rownode_dof.data(), sizeof(int32_t) * gh.rownode_dof.size()},
            {&p->d_nb_ptr, gh.nb_ptr.data(), sizeof(int32_t) * gh.nb_ptr.size()},
            {&p->d_nbcol, gh.nbcol.data(), sizeof(int32_t) * gh.nbcol.size()},
            {&p->d_diag_rank, gh.diag_rank.data(), sizeof(int32_t) * gh.diag_rank.size()},
            {&p->d_rn_meta, gh.rn_meta.data(), sizeof(int32_t) * gh.rn_meta.size()},
            {&p->d_ninc_ptr, gh.ninc_ptr.data(), sizeof(int32_t) * gh.ninc_ptr.size()},
            {&p->d_ninc, gh.ninc.data(), sizeof(int32_t) * gh.ninc.size()},
        };
        for (auto &u : rups) {
            if ((e = hipMalloc((void **)u.dst, u.bytes ? u.bytes : 4)) != hipSuccess) return fail(e, "hipMalloc");
            if (u.bytes && (e = hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice)) != hipSuccess)
                return fail(e, "hipMemcpy");
        }
    }
    if (p->npe == 8 || p->gather20) {
        Up gups[] = {
            {&p->d_gdesc, gh.gdesc.data(), sizeof(int32_t) * gh.gdesc.size()},
            {&p->d_gblob, gh.gblob.data(), sizeof(int32_t) * gh.gblob.size()},
            {&p->d_rorder, p->rorder.data(), sizeof(int32_t) * p->rorder.size()},

        };
        for (auto &u : gups) {
            if ((e = hipMalloc((void **)u.dst, u.bytes ? u.bytes : 4)) != hipSuccess) return fail(e, "hipMalloc");
            if (u.bytes && (e = hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice)) != hipSuccess)
                return fail(e, "hipMemcpy");
        }
    }
    if ((rc = alloc_records(p)) || (rc = ensure_pipeline(p))) {
        ewfe_plan_destroy(p);
        return rc;
    }
    p->use_gather20 = p->gather20;
    if (nnz_out) *nnz_out = p->nnz;
    *out = p;
    return 0;
}

extern "C" int ewfe_plan_set_row_gather(ewfe_plan *p, int32_t use_gather)
{
    if (!p) {
        ewfe::set_error("ewfe_plan_set_row_gather: null plan");
        return EWFE_EINVAL;
    }
    if (p->npe == 8) {
        if (use_gather) return 0;
        ewfe::set_error("ewfe_plan_set_row_gather: 8-node plans always assemble through the row-node gather");
        return EWFE_EINVAL;
    }
    if (use_gather && (!p->gather20 || !p->d_kbuf)) {
        ewfe::set_error("ewfe_plan_set_row_gather: the plan has no row-node gather (a node in more than 8 elements "
                        "or coupled with more than 128 nodes)");
        return EWFE_EINVAL;
    }
    p->use_gather20 = use_gather != 0;
    return 0;
}

extern "C" int ewfe_plan_create(ewfe_plan **out, int elem_kind, int64_t n_el, const int32_t *conn_host,
                                int64_t n_nodes, const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out)
{
    return plan_create(out, elem_kind, n_el, conn_host, n_nodes, node_dof_host, n_dof, nnz_out, true, 0);
}

extern "C" int ewfe_plan_create_chunked(ewfe_plan **out, int elem_kind, int64_t n_el, const int32_t *conn_host,
                                        int64_t n_nodes, const int64_t *node_dof_host, int64_t n_dof,
                                        int64_t *nnz_out, int pipeline_chunks)
{
    return plan_create(out, elem_kind, n_el, conn_host, n_nodes, node_dof_host, n_dof, nnz_out, true,
                       pipeline_chunks);
}

extern "C" int ewfe_plan_create_host(ewfe_plan **out, int elem_kind, int64_t n_el, const int32_t *conn_host,
                                     int64_t n_nodes, const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out,
                                     int pipeline_chunks)
{
    return plan_create(out, elem_kind, n_el, conn_host, n_nodes, node_dof_host, n_dof, nnz_out, false,
                       pipeline_chunks);
}

extern "C" int ewfe_plan_create_union(ewfe_plan **out, const ewfe_plan *const *plans, int32_t n_plans,
                                      int64_t n_nodes, const int64_t *node_dof_host, int64_t *nnz_out)
{
    if (!out || !plans || n_plans < 1 || n_nodes <= 0 || !node_dof_host) {
        ewfe::set_error("ewfe_plan_create_union: invalid arguments");
        return EWFE_EINVAL;
    }
    *out = nullptr;
    for (int k = 0; k < n_plans; ++k)
        if (!plans[k] || !plans[k]->d_indptr || plans[k]->n_dof != plans[0]->n_dof ||
            plans[k]->device != plans[0]->device || (int64_t)plans[k]->indptr.size() != plans[0]->n_dof + 1) {
            ewfe::set_error("ewfe_plan_create_union: every plan must be a device plan over the same dofs and device");
            return EWFE_EINVAL;
        }
    int dev;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev != plans[0]->device) {
        ewfe::set_error("ewfe_plan_create_union: current device differs from the plans' device");
        return EWFE_EINVAL;
    }
    ewfe_plan *u = new (std::nothrow) ewfe_plan();
    if (!u) {
        ewfe::set_error("out of host memory");
        return EWFE_ENOMEM;
    }
    u->elem_kind = -1;
    u->n_nodes = n_nodes;
    u->n_dof = plans[0]->n_dof;
    u->device = dev;
    ewfe::GatherHost gh;
    std::vector<std::vector<int32_t>> maps;
    int rc;
    try {
        rc = ewfe::build_union_host(*u, plans, n_plans, node_dof_host, n_nodes, gh, maps);
    } catch (const std::bad_alloc &) {
        rc = EWFE_ENOMEM;
        ewfe::set_error("out of host memory while building the system plan");
    }
    if (rc) {
        delete u;
        return rc;
    }
    auto up = [&](int32_t **dst, const void *src, size_t bytes) -> int {
        HIP_CHECK(hipMalloc((void **)dst, bytes ? bytes : 4));
        if (bytes) HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if ((rc = up(&u->d_indptr, u->indptr.data(), sizeof(int32_t) * u->indptr.size())) ||
        (rc = up(&u->d_indices, u->indices.data(), sizeof(int32_t) * u->indices.size())) ||
        (rc = up(&u->d_rownode_dof, gh.rownode_dof.data(), sizeof(int32_t) * gh.rownode_dof.size())) ||
        (rc = up(&u->d_nb_ptr, gh.nb_ptr.data(), sizeof(int32_t) * gh.nb_ptr.size())) ||
        (rc = up(&u->d_nbcol, gh.nbcol.data(), sizeof(int32_t) * gh.nbcol.size())) ||
        (rc = up(&u->d_diag_rank, gh.diag_rank.data(), sizeof(int32_t) * gh.diag_rank.size())) ||
        (rc = up(&u->d_rn_meta, gh.rn_meta.data(), sizeof(int32_t) * gh.rn_meta.size()))) {
        ewfe_plan_destroy(u);
        return rc;
    }
    u->d_umap.assign(n_plans, nullptr);
    u->umap_len.assign(n_plans, 0);
    for (int k = 0; k < n_plans; ++k) {
        if ((rc = up(&u->d_umap[k], maps[k].data(), sizeof(int32_t) * maps[k].size()))) {
            ewfe_plan_destroy(u);
            return rc;
        }
        u->umap_len[k] = (int64_t)maps[k].size();
    }
    if (nnz_out) *nnz_out = u->nnz;
    *out = u;
    return 0;
}

extern "C" int ewfe_plan_pipeline_orders(const ewfe_plan *p, int64_t *pipe_el_host, int64_t *pipe_row_host,
                                         int32_t *eorder_host, int32_t *rorder_host)
{
    if (!p || p->npe != 8) {
        ewfe::set_error("ewfe_plan_pipeline_orders: null plan or not an 8-node element plan");
        return EWFE_EINVAL;
    }
    if (pipe_el_host) std::memcpy(pipe_el_host, p->pipe_el.data(), sizeof(int64_t) * p->pipe_el.size());
    if (pipe_row_host) std::memcpy(pipe_row_host, p->pipe_row.data(), sizeof(int64_t) * p->pipe_row.size());
    if (eorder_host)
        for (int64_t i = 0; i < p->n_el; ++i) eorder_host[i] = (int32_t)p->order[i];
    if (rorder_host) std::memcpy(rorder_host, p->rorder.data(), sizeof(int32_t) * p->rorder.size());
    return 0;
}

extern "C" int ewfe_plan_pipeline_info(const ewfe_plan *p, int32_t *n_chunks, int32_t *structured)
{
    if (!p) {
        ewfe::set_error("ewfe_plan_pipeline_info: null plan");
        return EWFE_EINVAL;
    }
    if (n_chunks) *n_chunks = p->npe == 8 ? p->n_chunks : 0;
    if (structured) *structured = p->structured ? 1 : 0;
    return 0;
}

extern "C" int ewfe_plan_pattern(const ewfe_plan *p, int32_t *indptr_host, int32_t *indices_host)
{
    if (!p || !indptr_host || !indices_host) {
        ewfe::set_error("ewfe_plan_pattern: invalid arguments");
        return EWFE_EINVAL;
    }
    std::memcpy(indptr_host, p->indptr.data(), sizeof(int32_t) * (p->n_dof + 1));
    std::memcpy(indices_host, p->indices.data(), sizeof(int32_t) * p->nnz);
    return 0;
}

extern "C" int ewfe_plan_pattern_dev(const ewfe_plan *p, const int32_t **indptr_dev, const int32_t **indices_dev)
{
    if (!p || !indptr_dev || !indices_dev || !p->d_indptr) {
        ewfe::set_error("ewfe_plan_pattern_dev: invalid arguments or plan without device data");
        return EWFE_EINVAL;
    }
    *indptr_dev = p->d_indptr;
    *indices_dev = p->d_indices;
    return 0;
}

extern "C" int ewfe_plan_element_order(const ewfe_plan *p, int64_t *order_host, int32_t *n_colors_out)
{
    if (!p || !order_host) {
        ewfe::set_error("ewfe_plan_element_order: invalid arguments");
        return EWFE_EINVAL;
    }
    std::memcpy(order_host, p->order.data(), sizeof(int64_t) * p->n_el);
    if (n_colors_out) *n_colors_out = p->n_colors;
    return 0;
}

extern "C" int ewfe_plan_info(const ewfe_plan *p, int64_t *n_el, int32_t *npe, int32_t *nqp, int64_t *n_dof)
{
    if (!p) {
        ewfe::set_error("ewfe_plan_info: null plan");
        return EWFE_EINVAL;
    }
    if (n_el) *n_el = p->n_el;
    if (npe) *npe = p->npe;
    if (nqp) *nqp = p->nqp;
    if (n_dof) *n_dof = p->n_dof;
    return 0;
}

extern "C" int ewfe_plan_set_materials(ewfe_plan *p, int32_t n_mat, const double *params_host,
                                       const int32_t *material_of_element_host)
{
    if (!p || n_mat < 0 || (n_mat > 0 && (!params_host || !material_of_element_host))) {
        ewfe::set_error("ewfe_plan_set_materials: invalid arguments");
        return EWFE_EINVAL;
    }
    if (p->d_mtab) HIP_CHECK(hipFree(p->d_mtab));
    if (p->d_mid) HIP_CHECK(hipFree(p->d_mid));
    p->d_mtab = nullptr;
    p->d_mid = nullptr;
    p->n_mat = 0;
    if (n_mat == 0) return 0;
    std::vector<int32_t> mid(p->n_el);  // by plan position
    for (int64_t i = 0; i < p->n_el; ++i) {
        const int32_t m = material_of_element_host[p->order[i]];
        if (m < 0 || m >= n_mat) {
            ewfe::set_error("ewfe_plan_set_materials: material index outside [0, n_mat)");
            return EWFE_EINVAL;
        }
        mid[i] = m;
    }
    HIP_CHECK(hipMalloc((void **)&p->d_mtab, sizeof(double) * 8 * (size_t)n_mat));
    HIP_CHECK(hipMalloc((void **)&p->d_mid, sizeof(int32_t) * (size_t)std::max<int64_t>(p->n_el, 1)));
    HIP_CHECK(hipMemcpy(p->d_mtab, params_host, sizeof(double) * 8 * (size_t)n_mat, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(p->d_mid, mid.data(), sizeof(int32_t) * (size_t)p->n_el, hipMemcpyHostToDevice));
    p->n_mat = n_mat;
    return 0;
}

extern "C" int ewfe_plan_destroy(ewfe_plan *p)
{
    if (!p) return 0;
    void *bufs[] = {p->d_indptr,   p->d_indices, p->d_edof,    p->d_colpos,    p->d_order,
                    p->d_rownode_dof, p->d_gdesc, p->d_gblob, p->d_kbuf,
                    p->d_fe,       p->d_nb_ptr,  p->d_nbcol,   p->d_diag_rank, p->d_rn_meta,
                    p->d_rorder,   p->d_ninc_ptr, p->d_ninc,   p->d_bweights, p->d_fweights, p->d_pos_of_el,
                    p->d_fslot,    p->d_mtab,     p->d_mid};
    for (auto b : bufs)
        if (b) (void)hipFree(b);
    for (auto m : p->d_umap)
        if (m) (void)hipFree(m);
    for (void *ev : p->chunk_events)
        if (ev) (void)hipEventDestroy(reinterpret_cast<hipEvent_t>(ev));
    if (p->aux_stream) (void)hipStreamDestroy(reinterpret_cast<hipStream_t>(p->aux_stream));
    delete p;
    return 0;
}

// Element records of 8-node plans, allocated once when the plan is created (never inside an evaluation): sized
// for FP64 values and shared by both value precisions (FP32 records use the first half of each slot range).
static int alloc_records(ewfe_plan *p)
{
    if (p->d_kbuf) return 0;
    if (p->npe == 8) {
        HIP_CHECK(hipMalloc(&p->d_kbuf, sizeof(double) * ((size_t)kRec8 * p->n_el + kRecPad8)));
        HIP_CHECK(hipMalloc((void **)&p->d_fe, sizeof(double) * 24 * (size_t)p->n_el));
    } else if (p->gather20) {
        HIP_CHECK(hipMalloc(&p->d_kbuf, sizeof(double) * ((size_t)kRec20 * p->n_el + kRecPad20)));
        HIP_CHECK(hipMalloc((void **)&p->d_fe, sizeof(double) * 60 * (size_t)p->n_el));
    }
    return 0;
}

// The second stream and the chunk events of the pipelined 8-node evaluation (created with the plan, on the
// plan's device, default priority).
static int ensure_pipeline(ewfe_plan *p)
{
    if (p->n_chunks <= 1 || p->aux_stream) return 0;
    // events first: the stream is set last, so a failed set-up is retried in full by the next call
    p->chunk_events.resize(p->n_chunks + 1, nullptr);
    for (auto &ev : p->chunk_events) {
        if (ev) continue;
        hipEvent_t e;
        HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev = e;
    }
    hipStream_t st;
    HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    p->aux_stream = st;
    return 0;
}

// 8-node elements: chunk by chunk, the element pass writes the records of chunk c on the caller's stream
// and the gather of chunk c's row nodes runs on the plan's second stream once that pass is done, so the
// element pass of chunk c + 1 (arithmetic bound) overlaps the gather of chunk c (memory bound). The caller's
// stream waits for the last gather: seen from the caller, the call is ordered on its stream as before.
template <int FORM, int MAT, typename VT>
static int launch_hex8(ewfe_plan *p, EvalArgs A, hipStream_t s, bool mfma)
{
    if (!p->d_kbuf || !p->d_fe || (p->n_chunks > 1 && !p->aux_stream)) {
        ewfe::set_error("ewfe_eval_assemble: the plan has no element records");
        return EWFE_EINVAL;
    }
    const int K = p->n_chunks;
    hipStream_t g = K > 1 ? reinterpret_cast<hipStream_t>(p->aux_stream) : s;
    for (int c = 0; c < K; ++c) {
        RecordArgs R{p->d_kbuf, p->d_fe, p->pipe_el[c], p->pipe_el[c + 1], 0};
        const int64_t batches = (R.q_end - R.q_begin + 7) / 8;
        R.xcd_batches = (batches + 7) / 8;
        if (batches > 0) {
            const dim3 grid((unsigned)(8 * R.xcd_batches));
            // the matrix-core tangent exists for the materials of the rank form without the B term
            constexpr bool kHasMF = sizeof(VT) == 4 && MAT != M_SVK;
            if (kHasMF && mfma) {
                if (A.mtab)
                    hipLaunchKernelGGL((k_hex8_elem<FORM, MAT, VT, true, kHasMF>), grid, dim3(64), 0, s, A, R);
                else
                    hipLaunchKernelGGL((k_hex8_elem<FORM, MAT, VT, false, kHasMF>), grid, dim3(64), 0, s, A, R);
            } else if (A.mtab) {
                hipLaunchKernelGGL((k_hex8_elem<FORM, MAT, VT, true, false>), grid, dim3(64), 0, s, A, R);
            } else {
                hipLaunchKernelGGL((k_hex8_elem<FORM, MAT, VT, false, false>), grid, dim3(64), 0, s, A, R);
            }
            HIP_CHECK(hipGetLastError());
        }
        if (K > 1) {
            hipEvent_t ev = reinterpret_cast<hipEvent_t>(p->chunk_events[c]);
            HIP_CHECK(hipEventRecord(ev, s));
            HIP_CHECK(hipStreamWaitEvent(g, ev, 0));
        }
        const int64_t count = p->pipe_row[c + 1] - p->pipe_row[c];
        if (count <= 0) continue;
        const int64_t pairs = (count + 1) / 2;
        const int64_t chunk = (pairs + 7) / 8;
        if (p->structured) {
            GridArgs G{p->grid_el[0], p->grid_el[1], p->grid_el[2]};
            hipLaunchKernelGGL((k_gather8s<VT>), dim3((unsigned)(8 * chunk)), dim3(64), 0, g, G, p->d_rorder,
                               (const VT *)p->d_kbuf, (const double *)p->d_fe, (VT *)A.vals, A.P, A.F, A.d64,
                               p->pipe_row[c], count, chunk);
        } else {
            hipLaunchKernelGGL((k_gather8<VT>), dim3((unsigned)(8 * chunk)), dim3(64), 0, g,
                               reinterpret_cast<const int4 *>(p->d_gdesc), p->d_gblob, (const VT *)p->d_kbuf,
                               (const double *)p->d_fe, (VT *)A.vals, A.P, A.F, p->pipe_row[c], count, chunk);
        }
        HIP_CHECK(hipGetLastError());
    }
    if (K > 1) {
        hipEvent_t ev = reinterpret_cast<hipEvent_t>(p->chunk_events[K]);
        HIP_CHECK(hipEventRecord(ev, g));
        HIP_CHECK(hipStreamWaitEvent(s, ev, 0));
    }
    return 0;
}

// 20-node elements: with the row-node gather, one element pass over all positions writes the records and the
// gather sums them into every CSR value, P and F entry once; otherwise one launch per colour adds into the CSR
// values (no two elements of a colour share a node).
template <int FORM, int MAT, typename VT>
static int launch_hex20(ewfe_plan *p, EvalArgs A, hipStream_t s, bool mfma)
{
    constexpr int64_t kMaxGrid = (int64_t)1 << 24;  // workgroups per launch (64 lanes each)
    if (p->use_gather20) {
        if (!p->d_kbuf || !p->d_fe || !p->d_gdesc) {
            ewfe::set_error("ewfe_eval_assemble: the plan has no element records");
            return EWFE_EINVAL;
        }
        const RecordArgs R{p->d_kbuf, p->d_fe, 0, p->n_el, 0};
        for (int64_t b = 0; b < p->n_el; b += kMaxGrid) {
            A.p_begin = b;
            A.p_end = std::min<int64_t>(p->n_el, b + kMaxGrid);
            const dim3 grid((unsigned)(A.p_end - A.p_begin));
            // the matrix-core tangent: FP32 records of the rank-form materials
            constexpr bool kHasMF = sizeof(VT) == 4 && MAT != M_SVK;
            if (kHasMF && mfma) {
                if (p->nqp == 8)
                    hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 8, true, kHasMF>), grid, dim3(64), 0, s, A, R);
                else
                    hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 27, true, kHasMF>), grid, dim3(64), 0, s, A, R);
            } else if (p->nqp == 8) {
                hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 8, true>), grid, dim3(64), 0, s, A, R);
            } else {
                hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 27, true>), grid, dim3(64), 0, s, A, R);
            }
            HIP_CHECK(hipGetLastError());
        }
        for (int64_t b = 0; b < p->n_rownodes; b += kMaxGrid) {
            const int64_t count = std::min<int64_t>(p->n_rownodes - b, kMaxGrid);
            const int64_t chunk = (count + 7) / 8;
            hipLaunchKernelGGL((k_gather20<VT>), dim3((unsigned)(8 * chunk)), dim3(64), 0, s,
                               reinterpret_cast<const int4 *>(p->d_gdesc), p->d_gblob, (const VT *)p->d_kbuf,
                               (const double *)p->d_fe, (VT *)A.vals, A.P, A.F, b, count, chunk);
            HIP_CHECK(hipGetLastError());
        }
        return 0;
    }
    const RecordArgs R{nullptr, nullptr, 0, 0, 0};
    for (int c = 0; c < p->n_colors; ++c) {
        for (int64_t b = p->color_begin[c]; b < p->color_begin[c + 1]; b += kMaxGrid) {
            A.p_begin = b;
            A.p_end = std::min<int64_t>(p->color_begin[c + 1], b + kMaxGrid);
            const dim3 grid((unsigned)(A.p_end - A.p_begin));
            if (p->nqp == 8)
                hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 8, false>), grid, dim3(64), 0, s, A, R);
            else
                hipLaunchKernelGGL((k_hex20<FORM, MAT, VT, 27, false>), grid, dim3(64), 0, s, A, R);
            HIP_CHECK(hipGetLastError());
        }
    }
    return 0;
}

template <int FORM, int MAT, typename VT>
static int launch_family(ewfe_plan *p, EvalArgs A, hipStream_t s, bool mfma)
{
    if (p->npe == 8) return launch_hex8<FORM, MAT, VT>(p, A, s, mfma);
    return launch_hex20<FORM, MAT, VT>(p, A, s, mfma);
}

template <typename VT>
static int dispatch(ewfe_plan *p, int mat_kind, EvalArgs A, hipStream_t s, bool mfma)
{
    const bool tl = is_tl(p->elem_kind);
    if (!tl && mat_kind == EWFE_LINEAR_ELASTIC) return launch_family<FORM_SS, M_LE, VT>(p, A, s, mfma);
    if (!tl && mat_kind == EWFE_VON_MISES) return launch_family<FORM_SS, M_J2, VT>(p, A, s, mfma);
    if (tl && mat_kind == EWFE_NEO_HOOKE_WA) return launch_family<FORM_TL, M_NH, VT>(p, A, s, mfma);
    if (tl && mat_kind == EWFE_ST_VENANT_KIRCHHOFF) {
        if (mfma) {
            ewfe::set_error("ewfe_eval_assemble: the matrix-core FP32 tangent covers linear elastic, von Mises and "
                            "Neo-Hooke; St. Venant-Kirchhoff takes EWFE_F32 or EWFE_F64");
            return EWFE_EINVAL;
        }
        return launch_family<FORM_TL, M_SVK, VT>(p, A, s, false);
    }
    ewfe::set_error("unsupported element/material combination (small-strain elements take linear elastic or "
                    "von Mises, total-Lagrangian elements take Neo-Hooke or St. Venant-Kirchhoff)");
    return EWFE_EINVAL;
}

extern "C" int ewfe_eval_assemble(const ewfe_plan *p, int mat_kind, const double *mat_params_host,
                                  const double *el_param_dev, const double *X_dev, const double *U_np_dev,
                                  const double *dU_dev, const double *state_n_dev, double *state_t_dev, double *P_dev,
                                  double *F_dev, void *csr_vals_dev, int vals_dtype, double *pnewdt_dev, void *stream)
{
    if (!p || !mat_params_host || !X_dev || !U_np_dev || !dU_dev || !P_dev || !F_dev || !csr_vals_dev || !pnewdt_dev) {
        ewfe::set_error("ewfe_eval_assemble: null argument");
        return EWFE_EINVAL;
    }
    const int ns = ewfe_state_size(mat_kind);
    if (ns < 0) return ns;
    if (ns > 0 && (!state_n_dev || !state_t_dev)) {
        ewfe::set_error("ewfe_eval_assemble: material needs state_n/state_t buffers");
        return EWFE_EINVAL;
    }
    if (vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32 && vals_dtype != EWFE_F32_MFMA &&
        vals_dtype != EWFE_F32_D64 && vals_dtype != EWFE_F32_MFMA_D64) {
        ewfe::set_error("ewfe_eval_assemble: vals_dtype must be EWFE_F64, EWFE_F32, EWFE_F32_MFMA, EWFE_F32_D64 or "
                        "EWFE_F32_MFMA_D64");
        return EWFE_EINVAL;
    }
    const bool d64 = vals_dtype == EWFE_F32_D64 || vals_dtype == EWFE_F32_MFMA_D64;
    if (d64 && !ewfe::d64_supported(p)) {
        ewfe::set_error("ewfe_eval_assemble: EWFE_F32_D64 needs a plan whose row nodes cover every dof");
        return EWFE_EINVAL;
    }
    if (!p->d_edof) {
        ewfe::set_error("ewfe_eval_assemble: plan has no elements on the device (ewfe_plan_create_host, or a system "
                        "plan of ewfe_plan_create_union)");
        return EWFE_EINVAL;
    }
    int dev;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev != p->device) {
        ewfe::set_error("ewfe_eval_assemble: current device differs from the plan's device");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc;
    // the row-node gather writes P, F and every CSR value of the row nodes once; other dofs must read zero
    const bool gather = p->npe == 8 || p->use_gather20;
    if (!gather || 3 * p->n_rownodes != p->n_dof) {
        if ((rc = fill<double>(P_dev, p->n_dof, 0.0, s))) return rc;
        if ((rc = fill<double>(F_dev, p->n_dof, 0.0, s))) return rc;
    }
    if ((rc = fill<double>(pnewdt_dev, 1, 1e36, s))) return rc;
    if (!gather) {
        if (vals_dtype == EWFE_F64) {
            if ((rc = fill<double>(reinterpret_cast<double *>(csr_vals_dev), p->nnz, 0.0, s))) return rc;
        } else {
            if ((rc = fill<float>(reinterpret_cast<float *>(csr_vals_dev), p->nnz, 0.0f, s))) return rc;
        }
    }
    EvalArgs A;
    A.edof = p->d_edof;
    A.colpos = p->d_colpos;
    A.order = p->d_order;
    A.indptr = p->d_indptr;
    A.el_param = el_param_dev;
    A.mtab = p->n_mat ? p->d_mtab : nullptr;
    A.mid = p->d_mid;
    A.X = X_dev;
    A.U = U_np_dev;
    A.dU = dU_dev;
    A.state_n = state_n_dev;
    A.state_t = state_t_dev;
    A.P = P_dev;
    A.F = F_dev;
    A.vals = csr_vals_dev;
    A.d64 = d64 ? ewfe::d64_blocks(csr_vals_dev, p->nnz) : nullptr;
    A.pnewdt = pnewdt_dev;
    A.n_el = p->n_el;
    A.p_begin = A.p_end = 0;
    for (int i = 0; i < 8; ++i) A.mat.p[i] = mat_params_host[i];
    ewfe_plan *pm = const_cast<ewfe_plan *>(p);  // the pipelined evaluation creates its second stream once
    if (vals_dtype == EWFE_F64) return dispatch<double>(pm, mat_kind, A, s, false);
    const bool mfma = vals_dtype == EWFE_F32_MFMA || vals_dtype == EWFE_F32_MFMA_D64;
    if ((rc = dispatch<float>(pm, mat_kind, A, s, mfma))) return rc;
    // the structured 8-node gather writes the diagonal corrections itself; other paths take a pass of their own
    const bool fused = p->npe == 8 && p->structured;
    return d64 && !fused ? ewfe::rebuild_diag64(p, csr_vals_dev, stream) : 0;
}

extern "C" int ewfe_adjoint_param_grad(const ewfe_plan *p, int mat_kind, const double *mat_params_host,
                                       const double *el_param_dev, const double *X_dev, const double *U_dev,
                                       const double *lambda_dev, double *grad_dev, void *stream)
{
    if (!p || !mat_params_host || !X_dev || !U_dev || !lambda_dev || !grad_dev) {
        ewfe::set_error("ewfe_adjoint_param_grad: null argument");
        return EWFE_EINVAL;
    }
    if (!p->d_edof) {
        ewfe::set_error("ewfe_adjoint_param_grad: plan has no device data");
        return EWFE_EINVAL;
    }
    if (p->npe != 8) {
        ewfe::set_error("ewfe_adjoint_param_grad: 8-node elements only (C3D8, C3D8TL)");
        return EWFE_EINVAL;
    }
    const bool tl = is_tl(p->elem_kind);
    if (!((tl && mat_kind == EWFE_NEO_HOOKE_WA) || (!tl && mat_kind == EWFE_LINEAR_ELASTIC))) {
        ewfe::set_error("ewfe_adjoint_param_grad: supported are C3D8TL + Neo-Hooke (mu_e) and C3D8 + linear "
                        "elastic (E_e)");
        return EWFE_EINVAL;
    }
    EvalArgs A{};
    A.edof = p->d_edof;
    A.order = p->d_order;
    A.el_param = el_param_dev;
    A.mtab = p->n_mat ? p->d_mtab : nullptr;
    A.mid = p->d_mid;
    A.X = X_dev;
    A.U = U_dev;
    A.n_el = p->n_el;
    for (int i = 0; i < 8; ++i) A.mat.p[i] = mat_params_host[i];
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t blocks = (p->n_el + 7) / 8;
    if (tl)
        hipLaunchKernelGGL((k_adjoint8<FORM_TL, M_NH>), dim3((unsigned)blocks), dim3(64), 0, s, A, lambda_dev, grad_dev);
    else
        hipLaunchKernelGGL((k_adjoint8<FORM_SS, M_LE>), dim3((unsigned)blocks), dim3(64), 0, s, A, lambda_dev, grad_dev);
    HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename VT>
__global__ void k_csr_accumulate(const int32_t *__restrict__ map, int64_t n, const VT *__restrict__ src,
                                 VT *__restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[map[i]] += src[i];
}

extern "C" int ewfe_csr_accumulate(const ewfe_plan *u, int32_t block, const void *src_vals_dev, void *dst_vals_dev,
                                   int vals_dtype, void *stream)
{
    if (!u || block < 0 || block >= (int32_t)u->d_umap.size() || !src_vals_dev || !dst_vals_dev ||
        (vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32 && vals_dtype != EWFE_F32_D64) ||
        (vals_dtype == EWFE_F32_D64 && !ewfe::d64_supported(u))) {
        ewfe::set_error("ewfe_csr_accumulate: invalid arguments (a system plan of ewfe_plan_create_union and one of "
                        "its block indices)");
        return EWFE_EINVAL;
    }
    const int64_t n = u->umap_len[block];
    if (n == 0) return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    if (vals_dtype == EWFE_F64)
        hipLaunchKernelGGL(k_csr_accumulate<double>, dim3(blocks), dim3(256), 0, s, u->d_umap[block], n,
                           (const double *)src_vals_dev, (double *)dst_vals_dev);
    else
        hipLaunchKernelGGL(k_csr_accumulate<float>, dim3(blocks), dim3(256), 0, s, u->d_umap[block], n,
                           (const float *)src_vals_dev, (float *)dst_vals_dev);
    HIP_CHECK(hipGetLastError());
    // EWFE_F32_D64: the system's diagonal blocks rebuilt from its off-diagonal values as they stand (the block's own
    // FP64 blocks are not read; after the last block the system's are those of the summed matrix)
    return vals_dtype == EWFE_F32_D64 ? ewfe::rebuild_diag64(u, dst_vals_dev, stream) : 0;
}

extern "C" int ewfe_dirichlet_rows(const ewfe_plan *p, const int32_t *rows_dev, int64_t n_rows, void *csr_vals_dev,
                                   int vals_dtype, void *stream)
{
    if (!p || (n_rows > 0 && !rows_dev) || !csr_vals_dev) {
        ewfe::set_error("ewfe_dirichlet_rows: invalid arguments");
        return EWFE_EINVAL;
    }
    if (n_rows <= 0) return 0;
    if (!p->d_indptr) {
        ewfe::set_error("ewfe_dirichlet_rows: plan has no device data");
        return EWFE_EINVAL;
    }
    if ((vals_dtype != EWFE_F64 && vals_dtype != EWFE_F32 && vals_dtype != EWFE_F32_D64) ||
        (vals_dtype == EWFE_F32_D64 && !ewfe::d64_supported(p))) {
        ewfe::set_error("ewfe_dirichlet_rows: vals_dtype must be EWFE_F64, EWFE_F32 or EWFE_F32_D64 (the latter on a "
                        "plan whose row nodes cover every dof)");
        return EWFE_EINVAL;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t blocks = (n_rows + 3) / 4;
    if (vals_dtype == EWFE_F64)
        hipLaunchKernelGGL(k_dirichlet_rows<double>, dim3((unsigned)blocks), dim3(256), 0, s, rows_dev, n_rows,
                           p->d_indptr, p->d_indices, reinterpret_cast<double *>(csr_vals_dev));
    else
        hipLaunchKernelGGL(k_dirichlet_rows<float>, dim3((unsigned)blocks), dim3(256), 0, s, rows_dev, n_rows,
                           p->d_indptr, p->d_indices, reinterpret_cast<float *>(csr_vals_dev));
    HIP_CHECK(hipGetLastError());
    return vals_dtype == EWFE_F32_D64 ? ewfe::dirichlet_diag64(p, rows_dev, n_rows, csr_vals_dev, stream) : 0;
}

extern "C" int ewfe_body_force(ewfe_plan *p, const double *X_dev, const uint8_t *el_mask_dev,
                               const double *force_host, double *PExt_dev, void *stream)
{
    if (!p || !X_dev || !force_host || !PExt_dev) {
        ewfe::set_error("ewfe_body_force: null argument");
        return EWFE_EINVAL;
    }
    if (!p->d_edof || !p->d_ninc) {
        ewfe::set_error("ewfe_body_force: plan has no device data");
        return EWFE_EINVAL;
    }
    ewfe_plan *pm = p;  // the scratch weights are allocated by the plan on first use
    if (!pm->d_bweights) HIP_CHECK(hipMalloc((void **)&pm->d_bweights, sizeof(double) * (size_t)p->npe * p->n_el));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const unsigned eb = (unsigned)((p->n_el + 255) / 256);
    if (p->npe == 8)
        hipLaunchKernelGGL((k_body_weights<8, 8>), dim3(eb), dim3(256), 0, s, p->d_edof, p->d_order, X_dev, el_mask_dev,
                           pm->d_bweights, p->n_el);
    else if (p->nqp == 8)
        hipLaunchKernelGGL((k_body_weights<20, 8>), dim3(eb), dim3(256), 0, s, p->d_edof, p->d_order, X_dev,
                           el_mask_dev, pm->d_bweights, p->n_el);
    else
        hipLaunchKernelGGL((k_body_weights<20, 27>), dim3(eb), dim3(256), 0, s, p->d_edof, p->d_order, X_dev,
                           el_mask_dev, pm->d_bweights, p->n_el);
    HIP_CHECK(hipGetLastError());
    const unsigned nb = (unsigned)((p->n_rownodes + 255) / 256);
    hipLaunchKernelGGL(k_body_nodes, dim3(nb), dim3(256), 0, s, p->d_rownode_dof, p->d_ninc_ptr, p->d_ninc,
                       pm->d_bweights, p->npe, force_host[0], force_host[1], force_host[2], PExt_dev, p->n_rownodes);
    HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int ewfe_qp_results(const ewfe_plan *p, int mat_kind, const double *mat_params_host,
                               const double *el_param_dev, const double *X_dev, const double *U_dev,
                               const double *state_dev, int result, double *out_dev, void *stream)
{
    if (!p || !mat_params_host || !X_dev || !U_dev || !out_dev || (result != 0 && result != 1)) {
        ewfe::set_error("ewfe_qp_results: invalid arguments (result: 0 = stress, 1 = strain)");
        return EWFE_EINVAL;
    }
    if (!p->d_edof) {
        ewfe::set_error("ewfe_qp_results: plan has no device data");
        return EWFE_EINVAL;
    }
    const bool tl = is_tl(p->elem_kind);
    const bool ok = (!tl && (mat_kind == EWFE_LINEAR_ELASTIC || mat_kind == EWFE_VON_MISES)) ||
                    (tl && (mat_kind == EWFE_NEO_HOOKE_WA || mat_kind == EWFE_ST_VENANT_KIRCHHOFF));
    if (!ok) {
        ewfe::set_error("ewfe_qp_results: unsupported element/material combination");
        return EWFE_EINVAL;
    }
    if (mat_kind == EWFE_VON_MISES && result == 0 && !state_dev) {
        ewfe::set_error("ewfe_qp_results: the J2 stress needs the state");
        return EWFE_EINVAL;
    }
    EvalArgs A{};
    A.edof = p->d_edof;
    A.order = p->d_order;
    A.el_param = el_param_dev;
    A.mtab = p->n_mat ? p->d_mtab : nullptr;
    A.mid = p->d_mid;
    A.X = X_dev;
    A.U = U_dev;
    A.n_el = p->n_el;
    for (int i = 0; i < 8; ++i) A.mat.p[i] = mat_params_host[i];
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t nt = p->n_el * p->nqp;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
#define EWFE_QPR(NN, NQ, FORM, MAT) \
    hipLaunchKernelGGL((k_qp_results<NN, NQ, FORM, MAT>), dim3(blocks), dim3(256), 0, s, A, state_dev, result, out_dev)
    const bool svk = mat_kind == EWFE_ST_VENANT_KIRCHHOFF;
    if (p->npe == 8) {
        if (tl && svk) EWFE_QPR(8, 8, FORM_TL, M_SVK);
        else if (tl) EWFE_QPR(8, 8, FORM_TL, M_NH);
        else if (mat_kind == EWFE_VON_MISES) EWFE_QPR(8, 8, FORM_SS, M_J2);
        else EWFE_QPR(8, 8, FORM_SS, M_LE);
    } else if (p->nqp == 8) {
        if (tl && svk) EWFE_QPR(20, 8, FORM_TL, M_SVK);
        else if (tl) EWFE_QPR(20, 8, FORM_TL, M_NH);
        else if (mat_kind == EWFE_VON_MISES) EWFE_QPR(20, 8, FORM_SS, M_J2);
        else EWFE_QPR(20, 8, FORM_SS, M_LE);
    } else {
        if (tl && svk) EWFE_QPR(20, 27, FORM_TL, M_SVK);
        else if (tl) EWFE_QPR(20, 27, FORM_TL, M_NH);
        else if (mat_kind == EWFE_VON_MISES) EWFE_QPR(20, 27, FORM_SS, M_J2);
        else EWFE_QPR(20, 27, FORM_SS, M_LE);
    }
#undef EWFE_QPR
    HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int ewfe_surface_pressure(ewfe_plan *p, const double *X_dev, const int32_t *elements_dev,
                                     int64_t n_elements, int face, double pressure, double *PExt_dev, void *stream)
{
    if (!p || !X_dev || !PExt_dev || n_elements < 0 || (n_elements > 0 && !elements_dev) || face < 1 || face > 6) {
        ewfe::set_error("ewfe_surface_pressure: invalid arguments (face: 1..6)");
        return EWFE_EINVAL;
    }
    if (!p->d_edof || !p->d_ninc) {
        ewfe::set_error("ewfe_surface_pressure: plan has no device data");
        return EWFE_EINVAL;
    }
    if (n_elements == 0) return 0;
    // scratch, allocated by the plan on first use: element -> plan position, element -> list slot, and the
    // nodal forces of the listed elements (grown to the largest list seen)
    if (!p->d_pos_of_el) {
        std::vector<int32_t> inv(p->n_el);
        for (int64_t i = 0; i < p->n_el; ++i) inv[p->order[i]] = (int32_t)i;
        HIP_CHECK(hipMalloc((void **)&p->d_pos_of_el, sizeof(int32_t) * p->n_el));
        HIP_CHECK(hipMemcpy(p->d_pos_of_el, inv.data(), sizeof(int32_t) * p->n_el, hipMemcpyHostToDevice));
    }
    if (!p->d_fslot) HIP_CHECK(hipMalloc((void **)&p->d_fslot, sizeof(int32_t) * p->n_el));
    if (n_elements > p->fweights_cap) {
        if (p->d_fweights) HIP_CHECK(hipFree(p->d_fweights));
        p->d_fweights = nullptr;
        p->fweights_cap = 0;
        HIP_CHECK(hipMalloc((void **)&p->d_fweights, sizeof(double) * 3 * (size_t)p->npe * n_elements));
        p->fweights_cap = n_elements;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if ((rc = fill<int32_t>(p->d_fslot, p->n_el, -1, s))) return rc;
    const unsigned lb = (unsigned)((n_elements + 255) / 256);
    if (p->npe == 8)
        hipLaunchKernelGGL(k_face_pressure<8>, dim3(lb), dim3(256), 0, s, p->d_edof, p->d_pos_of_el, X_dev,
                           elements_dev, n_elements, face, pressure, p->d_fweights, p->d_fslot, p->n_el);
    else
        hipLaunchKernelGGL(k_face_pressure<20>, dim3(lb), dim3(256), 0, s, p->d_edof, p->d_pos_of_el, X_dev,
                           elements_dev, n_elements, face, pressure, p->d_fweights, p->d_fslot, p->n_el);
    HIP_CHECK(hipGetLastError());
    const unsigned nb = (unsigned)((p->n_rownodes + 255) / 256);
    hipLaunchKernelGGL(k_node_vec_sum, dim3(nb), dim3(256), 0, s, p->d_rownode_dof, p->d_ninc_ptr, p->d_ninc,
                       p->d_fweights, p->d_fslot, p->npe, PExt_dev, p->n_rownodes);
    HIP_CHECK(hipGetLastError());
    return 0;
}
