/*
 * ewfe_oracle.c -- CPU restatement of EdelweissFE's element-evaluation and
 * assembly path, used ONLY as test infrastructure (the checker in tests/, the
 * smoke check and the CPU-baseline leg of bench.py). Product code never links
 * or calls it.
 *
 * It follows the reference's in-tree formulation, step for step:
 *   - small-strain hexahedra: ew/elements/displacementelement/element.py:290-346
 *     (dE = B.dU; material update from the stored stress; K += B^T C B detJ w;
 *      P -= B^T sigma detJ w; strain += dE), state per quadrature point =
 *      stress(6) + strain(6) + material state (element.py:222-235);
 *   - VonMises J2 radial return with isotropic hardening and its consistent
 *     tangent: ew/materials/vonmises/vonmises.py:186-254 (constants :35-56,
 *     hardening :105-124);
 *   - isotropic linear elasticity: ew/materials/linearelastic/linearelastic.py:95-117,185-210;
 *   - total-Lagrangian hexahedra with a hyperelastic material:
 *     ew/elements/displacementtlelement/element.py:346-427 and
 *     _elementcomputationmatrices.py:70-107,247-306;
 *   - compressible Neo-Hooke (Pence-Gou formulation "a"):
 *     ew/materials/neohooke/neohookepencegouformulationa.py:107-141;
 *   - St. Venant-Kirchhoff (the Marmot material "StVenantKirchhoffIsotropic" of the
 *     C3D8TL regression deck, testfiles/C3D8TL/test.inp; Marmot is not in the
 *     reference tree, so the law is restated from its definition: S = lambda
 *     tr(E) I + 2 mu E with lambda, mu from E, nu; tau = F S F^T; its
 *     derivative dtau/dF in the element's formula below);
 *   - the parallel element loop (private Ke/Pe slots, then a serial P/F scatter
 *     in element order): ew/solvers/nonlinearimplicitstaticparallelmk2.pyx:111-198;
 *   - the VIJ -> CSR scatter-add in VIJ order: ew/numerics/csrgenerator.pyx:100-115.
 *
 * Shape functions are the textbook trilinear (8-node) and serendipity (20-node)
 * functions in Abaqus node order, which is the order boxGen emits
 * (ew/generators/boxgen.py:171-299). Gauss rules: 2x2x2 (w=1) and 3x3x3
 * (points 0, +-sqrt(0.6), weights 5/9, 8/9), ew/elements/library.py:33-47,212-275.
 * Quadrature points are in the reference's order (see gauss_point).
 *
 * Ke is written into the VIJ slice exactly as the reference element writes it:
 * reshape(K_, (nDof, nDof)) row-major, i.e. V[off + a*n + b] = Ke[a][b]
 * (element.py:318); the VIJ pattern maps that slot to (row idcs[b], col idcs[a])
 * (dofmanager.py:543-555).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

enum { EL_C3D8 = 0, EL_C3D20 = 1, EL_C3D8TL = 2, EL_C3D20TL = 3, EL_C3D20R = 4, EL_C3D20RTL = 5 };
enum { MAT_LE = 0, MAT_J2 = 1, MAT_NH = 2, MAT_SVK = 3 };

#define MAXN 20
#define MAXQ 27
#define MAXD 60

static int el_nnodes(int k) { return (k == EL_C3D8 || k == EL_C3D8TL) ? 8 : 20; }
static int el_nqp(int k) { return (k == EL_C3D20 || k == EL_C3D20TL) ? 27 : 8; } /* C3D20R(TL): 2x2x2 */
static int el_is_tl(int k) { return k == EL_C3D8TL || k == EL_C3D20TL || k == EL_C3D20RTL; }

/* Abaqus parametric corner/midside coordinates */
static const double HEX20_XI[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {0, -1, 1},  {1, 0, 1},  {0, 1, 1}, {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0}};

static void shape_deriv(int nn, const double p[3], double dN[][3], double *N)
{
    for (int a = 0; a < nn; ++a) {
        const double xa = HEX20_XI[a][0], ya = HEX20_XI[a][1], za = HEX20_XI[a][2];
        const double x = p[0], y = p[1], z = p[2];
        if (nn == 8) {
            const double fx = 1 + xa * x, fy = 1 + ya * y, fz = 1 + za * z;
            if (N) N[a] = 0.125 * fx * fy * fz;
            dN[a][0] = 0.125 * xa * fy * fz;
            dN[a][1] = 0.125 * fx * ya * fz;
            dN[a][2] = 0.125 * fx * fy * za;
        } else if (a < 8) {
            const double fx = 1 + xa * x, fy = 1 + ya * y, fz = 1 + za * z;
            const double s = xa * x + ya * y + za * z - 2;
            if (N) N[a] = 0.125 * fx * fy * fz * s;
            dN[a][0] = 0.125 * (xa * fy * fz * s + fx * fy * fz * xa);
            dN[a][1] = 0.125 * (fx * ya * fz * s + fx * fy * fz * ya);
            dN[a][2] = 0.125 * (fx * fy * za * s + fx * fy * fz * za);
        } else if (xa == 0) {
            const double gx = 1 - x * x, fy = 1 + ya * y, fz = 1 + za * z;
            if (N) N[a] = 0.25 * gx * fy * fz;
            dN[a][0] = 0.25 * (-2 * x) * fy * fz;
            dN[a][1] = 0.25 * gx * ya * fz;
            dN[a][2] = 0.25 * gx * fy * za;
        } else if (ya == 0) {
            const double fx = 1 + xa * x, gy = 1 - y * y, fz = 1 + za * z;
            if (N) N[a] = 0.25 * fx * gy * fz;
            dN[a][0] = 0.25 * xa * gy * fz;
            dN[a][1] = 0.25 * fx * (-2 * y) * fz;
            dN[a][2] = 0.25 * fx * gy * za;
        } else {
            const double fx = 1 + xa * x, fy = 1 + ya * y, gz = 1 - z * z;
            if (N) N[a] = 0.25 * fx * fy * gz;
            dN[a][0] = 0.25 * xa * fy * gz;
            dN[a][1] = 0.25 * fx * ya * gz;
            dN[a][2] = 0.25 * fx * fy * (-2 * z);
        }
    }
}

/* Quadrature point q in the reference's order (library.py:212-227, 260-275): 2x2x2: xi = -g for q < 4, +g
 * after, (eta, zeta) = g (-1,-1), (-1,+1), (+1,+1), (+1,-1) by q mod 4; 3x3x3: q = 9 i + 3 j + k over
 * (xi_i, eta_j, zeta_k). */
static void gauss_point(int nqp, int q, double p[3], double *w)
{
    if (nqp == 8) {
        const double g = 1.0 / sqrt(3.0);
        static const int eta[4] = {-1, -1, 1, 1}, zeta[4] = {-1, 1, 1, -1};
        p[0] = q < 4 ? -g : g;
        p[1] = eta[q & 3] * g;
        p[2] = zeta[q & 3] * g;
        *w = 1.0;
    } else {
        const double g = sqrt(0.6);
        const double pts[3] = {-g, 0.0, g};
        const double wts[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
        const int i = q / 9, j = (q / 3) % 3, k = q % 3;
        p[0] = pts[i];
        p[1] = pts[j];
        p[2] = pts[k];
        *w = wts[i] * wts[j] * wts[k];
    }
}

static double det3(const double A[3][3])
{
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

static double inv3(const double A[3][3], double B[3][3])
{
    const double d = det3(A);
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / d;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / d;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / d;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / d;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / d;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / d;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / d;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / d;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / d;
    return d;
}

/* Reference-frame gradients at one quadrature point:
 * J_ij = dx_j/dxi_i (rows = parametric directions, as _J3D8 builds it),
 * g_a = J^{-1} dN_a/dxi. Returns detJ. */
static double qp_gradients(int nn, int nqp, int q, const double *X, double g[][3], double *w)
{
    double p[3], dN[MAXN][3], J[3][3] = {{0}}, Ji[3][3];
    gauss_point(nqp, q, p, w);
    shape_deriv(nn, p, dN, NULL);
    for (int a = 0; a < nn; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * X[3 * a + j];
    const double d = inv3(J, Ji);
    for (int a = 0; a < nn; ++a)
        for (int j = 0; j < 3; ++j) g[a][j] = Ji[j][0] * dN[a][0] + Ji[j][1] * dN[a][1] + Ji[j][2] * dN[a][2];
    return d;
}

/* 3D isotropic elasticity matrix, Voigt [11,22,33,12,13,23], engineering shear
 * (linearelastic.py:95-117). */
static void elasticity(double E, double nu, double C[6][6])
{
    memset(C, 0, 36 * sizeof(double));
    const double f = E / ((1 + nu) * (1 - 2 * nu));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i][j] = f * (i == j ? (1 - nu) : nu);
    for (int i = 3; i < 6; ++i) C[i][i] = f * (1 - 2 * nu) / 2;
}

/* VonMisesMaterial.computeStress (vonmises.py:186-254). kappa is the single
 * material state variable. Returns 0, or 1 for a cutback request (0.5). */
static int j2_compute_stress(const double *mp, double *stress, double C[6][6], const double *dE, double *kappa)
{
    const double E = mp[0], nu = mp[1], fy0 = mp[2], HLin = mp[3], dfy = mp[4], delta = mp[5];
    const double G = E / (2 * (1.0 + nu));
    double Ei[6][6];
    elasticity(E, nu, Ei);
    double nrm = 0;
    for (int i = 0; i < 6; ++i) nrm += dE[i] * dE[i];
    nrm = sqrt(nrm);
    if (nrm < 1e-14) {
        memcpy(C, Ei, sizeof(Ei));
        return 0;
    }
    double tr[6];
    for (int j = 0; j < 6; ++j) {
        double s = 0;
        for (int i = 0; i < 6; ++i) s += dE[i] * Ei[i][j];
        tr[j] = stress[j] + s;
    }
    /* deviatoric(): IDev @ stress, IDev has 1 on the shear diagonal */
    const double p3 = (tr[0] + tr[1] + tr[2]) / 3.0;
    double dev[6];
    for (int i = 0; i < 3; ++i) dev[i] = tr[i] - p3;
    for (int i = 3; i < 6; ++i) dev[i] = tr[i];
    const double devNorm = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                                2 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    const double k0 = kappa[0];
#define FY(k) (fy0 + HLin * (k) + dfy * (1.0 - exp(-delta * (k))))
#define DFY(k) (HLin + dfy * delta * exp(-delta * (k)))
    if (devNorm - sqrt(2.0 / 3.0) * FY(k0) > 0.0) {
        double n[6];
        for (int i = 0; i < 6; ++i) n[i] = dev[i] / devNorm;
        int counter = 0;
        double dK = 0;
        for (;;) {
            const double R = devNorm - sqrt(6.0) * G * dK - sqrt(2.0 / 3.0) * FY(k0 + dK);
            if (!(fabs(R) > 1e-12)) break;
            if (counter == 15) return 1;
            const double dR = -sqrt(6.0) * G - sqrt(2.0 / 3.0) * DFY(k0 + dK);
            dK -= R / dR;
            counter++;
        }
        const double dLambda = sqrt(3.0 / 2.0) * dK;
        kappa[0] = k0 + dK;
        for (int i = 0; i < 6; ++i) stress[i] = tr[i] - 2.0 * G * dLambda * n[i];
        const double a1 = 2.0 * G * (1.0 / (1.0 + DFY(kappa[0]) / (3.0 * G)) - 2.0 * G * dLambda / devNorm);
        const double a2 = 4.0 * G * G * dLambda / devNorm;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                double idev;
                if (i < 3 && j < 3)
                    idev = (i == j) ? 2.0 / 3.0 : -1.0 / 3.0;
                else
                    idev = (i == j) ? 0.5 : 0.0;
                C[i][j] = Ei[i][j] - a1 * n[i] * n[j] - a2 * idev;
            }
    } else {
        memcpy(stress, tr, sizeof(tr));
        memcpy(C, Ei, sizeof(Ei));
    }
#undef FY
#undef DFY
    return 0;
}

/* Small-strain hexahedron (DisplacementElement.computeYourself). state layout
 * per QP: [stress 6 | strain 6 | material state]. Ke row-major [a][b]. */
static int small_strain_element(int ek, int mk, const double *mp, const double *X, const double *dU,
                                const double *state_n, double *state_t, int nstate, double *Ke, double *Pe)
{
    const int nn = el_nnodes(ek), nq = el_nqp(ek), nd = 3 * nn;
    memset(Ke, 0, sizeof(double) * nd * nd);
    memset(Pe, 0, sizeof(double) * nd);
    memcpy(state_t, state_n, sizeof(double) * nq * nstate);
    for (int q = 0; q < nq; ++q) {
        double g[MAXN][3], w;
        const double detJ = qp_gradients(nn, nq, q, X, g, &w);
        /* B (6 x nd), rows [11,22,33,12,13,23], engineering shear */
        double B[6][MAXD];
        memset(B, 0, sizeof(B));
        for (int a = 0; a < nn; ++a) {
            const double gx = g[a][0], gy = g[a][1], gz = g[a][2];
            B[0][3 * a + 0] = gx;
            B[1][3 * a + 1] = gy;
            B[2][3 * a + 2] = gz;
            B[3][3 * a + 0] = gy;
            B[3][3 * a + 1] = gx;
            B[4][3 * a + 0] = gz;
            B[4][3 * a + 2] = gx;
            B[5][3 * a + 1] = gz;
            B[5][3 * a + 2] = gy;
        }
        double dE[6];
        for (int i = 0; i < 6; ++i) {
            double s = 0;
            for (int k = 0; k < nd; ++k) s += B[i][k] * dU[k];
            dE[i] = s;
        }
        double *st = state_t + q * nstate;
        double C[6][6];
        if (mk == MAT_LE) {
            elasticity(mp[0], mp[1], C);
            for (int j = 0; j < 6; ++j) {
                double s = 0;
                for (int i = 0; i < 6; ++i) s += C[j][i] * dE[i];
                st[j] += s;
            }
        } else {
            if (j2_compute_stress(mp, st, C, dE, st + 12)) return 1;
        }
        const double f = detJ * w;
        double CB[6][MAXD];
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < nd; ++k) {
                double s = 0;
                for (int j = 0; j < 6; ++j) s += C[i][j] * B[j][k];
                CB[i][k] = s;
            }
        for (int a = 0; a < nd; ++a) {
            for (int b = 0; b < nd; ++b) {
                double s = 0;
                for (int i = 0; i < 6; ++i) s += B[i][a] * CB[i][b];
                Ke[a * nd + b] += s * f;
            }
            double s = 0;
            for (int i = 0; i < 6; ++i) s += B[i][a] * st[i];
            Pe[a] -= s * f;
        }
        for (int i = 0; i < 6; ++i) st[6 + i] += dE[i];
    }
    return 0;
}

/* Total-Lagrangian hexahedron with a hyperelastic material (DisplacementTLElement,
 * hyperelastic branch). U is the total displacement. mk = MAT_NH: mp = [mu, K];
 * mk = MAT_SVK: mp = [E, nu]. state per QP: [Kirchhoff stress 6 (11,22,33,12,23,13) |
 * Green strain 6 | energy]. */
static int tl_element(int ek, int mk, const double *mp, const double *X, const double *U, double *state_t,
                      int nstate, double *Ke, double *Pe)
{
    const int nn = el_nnodes(ek), nq = el_nqp(ek), nd = 3 * nn;
    const double mu = mp[0], Kb = mp[1];
    const double lamS = mp[0] * mp[1] / ((1 + mp[1]) * (1 - 2 * mp[1])), muS = mp[0] / (2 * (1 + mp[1]));
    memset(Ke, 0, sizeof(double) * nd * nd);
    memset(Pe, 0, sizeof(double) * nd);
    for (int q = 0; q < nq; ++q) {
        double g[MAXN][3], w;
        const double detJ = qp_gradients(nn, nq, q, X, g, &w);
        double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int a = 0; a < nn; ++a)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) F[i][j] += U[3 * a + i] * g[a][j];
        double Fi[3][3];
        const double J = inv3(F, Fi);
        double Bc[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Bc[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
        const double I1 = Bc[0][0] + Bc[1][1] + Bc[2][2];
        const double lamBar = (Kb - 2.0 / 3.0 * mu) * (J * J - J) - mu;
        const double muBar = (Kb - 2.0 / 3.0 * mu) * (2 * J * J - J);
        double T[3][3];
        double S2[3][3], Egr[3][3], trE = 0;  /* SVK: second Piola-Kirchhoff stress, Green strain */
        if (mk == MAT_SVK) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    Egr[i][j] = 0.5 * (F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j] - (i == j));
            trE = Egr[0][0] + Egr[1][1] + Egr[2][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) S2[i][j] = 2 * muS * Egr[i][j] + (i == j ? lamS * trE : 0.0);
            double FS[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) FS[i][j] = F[i][0] * S2[0][j] + F[i][1] * S2[1][j] + F[i][2] * S2[2][j];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) T[i][j] = FS[i][0] * F[j][0] + FS[i][1] * F[j][1] + FS[i][2] * F[j][2];
        } else {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) T[i][j] = mu * Bc[i][j] + (i == j ? lamBar : 0.0);
        }
        double *st = state_t + q * nstate;
        st[0] = T[0][0];
        st[1] = T[1][1];
        st[2] = T[2][2];
        st[3] = T[0][1];
        st[4] = T[1][2];
        st[5] = T[0][2];
        double H[3][3], Eg[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[i][j] = F[i][j] - (i == j);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Eg[i][j] = 0.5 * (H[i][j] + H[j][i] + H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
        st[6] = Eg[0][0];
        st[7] = Eg[1][1];
        st[8] = Eg[2][2];
        st[9] = 2 * Eg[0][1];
        st[10] = 2 * Eg[1][2];
        st[11] = 2 * Eg[0][2];
        if (nstate > 12) {
            if (mk == MAT_SVK) {
                double EE = 0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) EE += Egr[i][j] * Egr[i][j];
                st[12] = 0.5 * lamS * trE * trE + muS * EE;
            } else {
                st[12] = mu / 2 * (I1 - 3) + (Kb / 2 - mu / 3) * (J - 1) * (J - 1) - mu * log(J);
            }
        }
        /* A_ijkl = dtau_ij/dF_kl */
        double A[3][3][3][3];
        if (mk == MAT_SVK) {
            /* tau = F S F^T, dS_mn/dF_kl = lambda delta_mn F_kl + mu (delta_ml F_kn + F_km delta_nl):
             * A_ijkl = delta_ik (F S)_jl + (F S)_il delta_jk + lambda B_ij F_kl + mu (F_il B_kj + B_ik F_jl) */
            double FS[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) FS[i][j] = F[i][0] * S2[0][j] + F[i][1] * S2[1][j] + F[i][2] * S2[2][j];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 3; ++l)
                            A[i][j][k][l] = (i == k) * FS[j][l] + FS[i][l] * (j == k) + lamS * Bc[i][j] * F[k][l] +
                                            muS * (F[i][l] * Bc[k][j] + Bc[i][k] * F[j][l]);
        } else {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 3; ++l)
                            A[i][j][k][l] =
                                mu * ((i == k) * F[j][l] + F[i][l] * (j == k)) + muBar * (i == j) * Fi[l][k];
        }
        /* spatial gradients NAi[a][m] = sum_I g[a][I] Fi[I][m] */
        double NA[MAXN][3];
        for (int a = 0; a < nn; ++a)
            for (int m = 0; m < 3; ++m) NA[a][m] = g[a][0] * Fi[0][m] + g[a][1] * Fi[1][m] + g[a][2] * Fi[2][m];
        /* PK1^T = F^{-1} tau */
        double PK[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) PK[i][j] = Fi[i][0] * T[0][j] + Fi[i][1] * T[1][j] + Fi[i][2] * T[2][j];
        const double f = detJ * w;
        for (int a = 0; a < nn; ++a)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int i = 0; i < 3; ++i) s += g[a][i] * PK[i][j];
                Pe[3 * a + j] -= s * f;
            }
        for (int a = 0; a < nn; ++a)
            for (int b = 0; b < nn; ++b) {
                double NT[3]; /* sum_i NA[b][i] T[i][j] */
                for (int j = 0; j < 3; ++j) NT[j] = NA[b][0] * T[0][j] + NA[b][1] * T[1][j] + NA[b][2] * T[2][j];
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        double s = 0;
                        for (int i = 0; i < 3; ++i)
                            for (int l = 0; l < 3; ++l) s += NA[a][i] * A[i][j][k][l] * g[b][l];
                        s -= NA[a][k] * NT[j];
                        Ke[(3 * a + j) * nd + 3 * b + k] += s * f;
                    }
            }
    }
    return 0;
}

/* One element: the MarmotElement::computeYourself contract
 * (ew/elements/marmotelement/element.pxd:97-103): Ke, Pe out, pNewDT < 1
 * requests a cutback. state_t receives a copy of state_n, then the update. */
int oracle_element(int ek, int mk, const double *mp, const double *X, const double *U, const double *dU,
                   const double *state_n, double *state_t, int nstate, double *Ke, double *Pe, double *pNewDT)
{
    int rc;
    if (el_is_tl(ek)) {
        memcpy(state_t, state_n, sizeof(double) * el_nqp(ek) * nstate);
        rc = tl_element(ek, mk, mp, X, U, state_t, nstate, Ke, Pe);
    } else {
        rc = small_strain_element(ek, mk, mp, X, dU, state_n, state_t, nstate, Ke, Pe);
    }
    if (rc) *pNewDT = 0.5;
    return rc;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/*
 * NISTParallel.computeElements + CSRGenerator.updateCSR over an element list.
 *  conn      int32 [n_el * nn] node indices (element-local order)
 *  X         f64 [n_nodes * 3]
 *  edofs     int64 [n_el * nd] element dof index list (dofmanager.py:445-471)
 *  el_mp8    optional f64 [n_el * 8] per-element parameter sets (several materials, each element its
 *            section's material, sectionbase.py:84-112), replacing mp; or NULL
 *  el_mp     optional per-element override of mp[0] (mu_e / E_e), or NULL
 *  V         f64 [n_el * nd * nd] VIJ values
 *  x_slot    int32 [n_el * nd * nd] VIJ position -> CSR slot (csrgenerator.pyx:78-97), or NULL
 *  csr_data  f64 [nnz] (zeroed here, then filled in VIJ order), or NULL
 *  times[2]  out: seconds in "elements" (loop + scatter) and "CSR generation"
 * Returns 0, or the (1-based) index of the first element that requested a cutback.
 */
int64_t oracle_eval_assemble(int ek, int mk, const double *mp, const double *el_mp8, const double *el_mp, int64_t n_el,
                             const int32_t *conn, const double *X, const int64_t *edofs, const double *U_np,
                             const double *dU, const double *state_n, double *state_t, int nstate, double *P,
                             double *F, double *V, const int32_t *x_slot, int64_t nnz, double *csr_data,
                             int n_threads, double *pNewDT, double *times)
{
    const int nn = el_nnodes(ek), nq = el_nqp(ek), nd = 3 * nn;
    double *Pe_all = (double *)malloc(sizeof(double) * n_el * nd);
    int64_t first_fail = 0;
    double t0 = now_s();
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t e = 0; e < n_el; ++e) {
        double Xe[3 * MAXN], Ue[MAXD], dUe[MAXD], mpe[8];
        for (int a = 0; a < nn; ++a)
            for (int c = 0; c < 3; ++c) Xe[3 * a + c] = X[3 * (int64_t)conn[e * nn + a] + c];
        for (int k = 0; k < nd; ++k) {
            Ue[k] = U_np[edofs[e * nd + k]];
            dUe[k] = dU[edofs[e * nd + k]];
        }
        for (int i = 0; i < 8; ++i) mpe[i] = el_mp8 ? el_mp8[e * 8 + i] : mp[i];
        if (el_mp) mpe[0] = el_mp[e];
        double pdt = 1e36;
        oracle_element(ek, mk, mpe, Xe, Ue, dUe, state_n + e * nq * nstate, state_t + e * nq * nstate, nstate,
                       V + e * nd * nd, Pe_all + e * nd, &pdt);
        if (pdt < 1.0) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                if (pdt < *pNewDT) *pNewDT = pdt;
                if (!first_fail || e + 1 < first_fail) first_fail = e + 1;
            }
        }
    }
    /* serial scatter in element order (mk2.pyx:186-193) */
    for (int64_t e = 0; e < n_el; ++e)
        for (int k = 0; k < nd; ++k) {
            const int64_t d = edofs[e * nd + k];
            P[d] += Pe_all[e * nd + k];
            F[d] += fabs(Pe_all[e * nd + k]);
        }
    double t1 = now_s();
    if (csr_data && x_slot) {
        memset(csr_data, 0, sizeof(double) * nnz);
        const int64_t nv = n_el * nd * nd;
        for (int64_t k = 0; k < nv; ++k) csr_data[x_slot[k]] += V[k];
    }
    double t2 = now_s();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
    }
    free(Pe_all);
    return first_fail;
}

/* Body force, restating DisplacementElement.computeBodyForce (edelweissfe/elements/displacementelement/
 * element.py:348-371): P_e += sum_q outer(N(xi_q), b) detJ_q w_q (reference configuration), and its assembly
 * NIST.computeBodyForces (edelweissfe/solvers/nonlinearimplicitstatic.py:503-557): PExt[el] += Pe over the
 * load's element set, in element order. mask: element -> 0/1 (NULL: every element).
 * Pinned by the closed forms of a rectangular brick (V/8 per C3D8 node; -V/8 per C3D20 corner and V/6 per
 * mid-side node) and by the reference's WallShearHexa8 / WallShearHexa20 decks (U.ref). */
void oracle_body_force(int ek, int64_t n_el, const int32_t *conn, const double *X, const int64_t *edofs,
                       const uint8_t *mask, const double *b, double *PExt)
{
    const int nn = el_nnodes(ek), nq = el_nqp(ek), nd = 3 * nn;
    for (int64_t e = 0; e < n_el; ++e) {
        if (mask && !mask[e]) continue;
        double Xe[3 * MAXN], Pe[MAXD];
        for (int a = 0; a < nn; ++a)
            for (int c = 0; c < 3; ++c) Xe[3 * a + c] = X[3 * (int64_t)conn[e * nn + a] + c];
        for (int k = 0; k < nd; ++k) Pe[k] = 0.0;
        for (int q = 0; q < nq; ++q) {
            double p[3], w, dN[MAXN][3], N[MAXN], J[3][3] = {{0}};
            gauss_point(nq, q, p, &w);
            shape_deriv(nn, p, dN, N);
            for (int a = 0; a < nn; ++a)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * Xe[3 * a + j];
            const double detJ = det3(J);
            for (int a = 0; a < nn; ++a)
                for (int c = 0; c < 3; ++c) Pe[3 * a + c] += N[a] * b[c] * detJ * w;
        }
        for (int k = 0; k < nd; ++k) PExt[edofs[e * nd + k]] += Pe[k];
    }
}

/* Pressure on one face of the masked elements (the Marmot element's computeDistributedLoad for type pressure,
 * called by nonlinearimplicitstatic.py:460-501): Pe_a = -p * sum_g N_a (dX/du x dX/dv) s w_g on the face with
 * Abaqus number `face` (1: zeta=-1, 2: zeta=+1, 3: eta=-1, 4: xi=+1, 5: eta=+1, 6: xi=-1), (u, v) the parametric
 * axes after the fixed one in cyclic order and s the sign of the fixed coordinate, so that the normal points
 * out of the element; Gauss rule 2x2 (8 nodes) or 3x3 (20 nodes); reference configuration. */
void oracle_surface_pressure(int ek, int64_t n_el, const int32_t *conn, const double *X, const int64_t *edofs,
                             const uint8_t *mask, int face, double pressure, double *PExt)
{
    static const int axis[7] = {0, 2, 2, 1, 0, 1, 0}, sign[7] = {0, -1, 1, -1, 1, 1, -1};
    const int nn = el_nnodes(ek), nd = 3 * nn, ng = nn == 8 ? 2 : 3;
    const int ax = axis[face], u = (ax + 1) % 3, v = (ax + 2) % 3;
    const double s = sign[face];
    const double g2[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)}, w2[2] = {1.0, 1.0};
    const double g3[3] = {-sqrt(0.6), 0.0, sqrt(0.6)}, w3[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int64_t e = 0; e < n_el; ++e) {
        if (mask && !mask[e]) continue;
        double Xe[3 * MAXN], Pe[MAXD];
        for (int a = 0; a < nn; ++a)
            for (int c = 0; c < 3; ++c) Xe[3 * a + c] = X[3 * (int64_t)conn[e * nn + a] + c];
        for (int k = 0; k < nd; ++k) Pe[k] = 0.0;
        for (int i = 0; i < ng; ++i)
            for (int j = 0; j < ng; ++j) {
                double p[3], dN[MAXN][3], N[MAXN], tu[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
                p[ax] = s;
                p[u] = ng == 2 ? g2[i] : g3[i];
                p[v] = ng == 2 ? g2[j] : g3[j];
                const double w = ng == 2 ? w2[i] * w2[j] : w3[i] * w3[j];
                shape_deriv(nn, p, dN, N);
                for (int a = 0; a < nn; ++a)
                    for (int c = 0; c < 3; ++c) {
                        tu[c] += dN[a][u] * Xe[3 * a + c];
                        tv[c] += dN[a][v] * Xe[3 * a + c];
                    }
                const double n[3] = {s * (tu[1] * tv[2] - tu[2] * tv[1]), s * (tu[2] * tv[0] - tu[0] * tv[2]),
                                     s * (tu[0] * tv[1] - tu[1] * tv[0])};
                for (int a = 0; a < nn; ++a)
                    for (int c = 0; c < 3; ++c) Pe[3 * a + c] -= pressure * N[a] * n[c] * w;
            }
        for (int k = 0; k < nd; ++k) PExt[edofs[e * nd + k]] += Pe[k];
    }
}
