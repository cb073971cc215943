/*
 * ewfe.h -- C ABI of libewfe.so: MI355X (gfx950) element evaluation + CSR assembly for the
 * nonlinear implicit static Newton loop of EdelweissFE.
 *
 * Every entry point returns 0 on success and a negative code on failure; ewfe_last_error()
 * then holds a message. No C++ exception crosses this boundary. Buffers named *_dev are device
 * pointers (e.g. torch tensors' data_ptr()), owned by the caller; buffers named *_host are host
 * memory. A plan owns only its own index structures (element order, slot map, colouring).
 * Calls are ordered on the HIP stream passed in (NULL = default stream); distinct plans may be
 * used concurrently from different host threads.
 *
 * Reference interfaces these entry points replace (paths relative to the EdelweissFE tree):
 *   ewfe_plan_create / ewfe_plan_pattern
 *       DofManager VIJ pattern + CSRGenerator.__init__:
 *       edelweissfe/numerics/dofmanager.py:522-557, edelweissfe/numerics/csrgenerator.pyx:47-98
 *   ewfe_eval_assemble
 *       NIST.computeElements + NIST.assembleStiffnessCSR (CSRGenerator.updateCSR), i.e. the
 *       per-element MarmotElement::computeYourself(QTotal, dQ, Pe, Ke, time, dT, pNewDT) loop:
 *       edelweissfe/solvers/nonlinearimplicitstatic.py:794-849, 753-769;
 *       edelweissfe/solvers/nonlinearimplicitstaticparallelmk2.pyx:111-198;
 *       edelweissfe/elements/marmotelement/element.pxd:97-103; csrgenerator.pyx:100-115
 *   ewfe_dirichlet_rows
 *       NIST.applyDirichletK (row elimination + unit diagonal): nonlinearimplicitstatic.py:559-593
 *   ewfe_adjoint_param_grad
 *       no reference counterpart (adjoint of the per-element parameter field of
 *       edelweissfe/sections/base/sectionbase.py:86-128)
 *   ewfe_spmv / ewfe_pcg / ewfe_pcg_workspace_size
 *       NIST.linearSolve (nonlinearimplicitstatic.py:727-751) and the linear-solver registry
 *       (edelweissfe/config/linsolve.py:46-86)
 */
#ifndef EWFE_H
#define EWFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ewfe_plan ewfe_plan;
typedef struct ewfe_mg ewfe_mg;

enum ewfe_elem_kind {
    EWFE_C3D8 = 0,
    EWFE_C3D20 = 1,
    EWFE_C3D8TL = 2,
    EWFE_C3D20TL = 3,
    EWFE_C3D20R = 4,   /* 20 nodes, 2x2x2 points (edelweissfe/elements/library.py:276-290) */
    EWFE_C3D20RTL = 5  /* the same, total Lagrangian (library.py:517-531) */
};
/* Materials: small-strain elements take EWFE_LINEAR_ELASTIC or EWFE_VON_MISES, total-Lagrangian elements
 * EWFE_NEO_HOOKE_WA or EWFE_ST_VENANT_KIRCHHOFF (Marmot's StVenantKirchhoffIsotropic of the reference's C3D8TL
 * regression deck, testfiles/C3D8TL/test.inp: S = lambda tr(E) I + 2 mu E with E the Green strain; no
 * matrix-core tangent). */
enum ewfe_mat_kind {
    EWFE_LINEAR_ELASTIC = 0,
    EWFE_VON_MISES = 1,
    EWFE_NEO_HOOKE_WA = 2,
    EWFE_ST_VENANT_KIRCHHOFF = 3
};
/* Precision of the assembled tangent. EWFE_F32 and EWFE_F32_MFMA both store FP32 CSR values. EWFE_F32 forms the
 * element tangents of 8-node elements in FP32 arithmetic on the vector units (per-point data rounded to FP32); those
 * of 20-node elements are formed with FP64 FMAs and rounded when they are stored; EWFE_F32_MFMA forms them in FP32 on
 * the matrix cores (v_mfma_f32_16x16x4_f32) for 8-node elements and for 20-node elements assembled through the
 * row-node gather (20-node plans on the colour path take it as EWFE_F32). The residual, nodal forces and state stay FP64. Only ewfe_eval_assemble takes EWFE_F32_MFMA; every other
 * call takes EWFE_F32 for FP32 values.
 * EWFE_F32_D64 stores the EWFE_F32 values (formed the same way, bitwise equal) followed, at byte offset
 * ewfe_vals_bytes(plan, EWFE_F32) rounded up to 256, by FP64 corrections of the diagonal blocks: 9 doubles per row
 * node (row-major, node of first dof 3 r at block r), E = -(sum of the node rows' off-diagonal FP32 blocks) - (its
 * FP32 diagonal block). The matrix these values stand for has the diagonal blocks FP32 block + E, i.e. minus the
 * sums of the off-diagonal FP32 blocks: an element's tangent maps a rigid translation to zero (its shape-function
 * gradients sum to zero); rounding the values to FP32 breaks that, the corrections restore it exactly, which keeps
 * Newton's contraction with an FP32 tangent close to that of the FP64 one. Every product, residual, Dirichlet edit
 * and CG iteration adds them; the multigrid hierarchy is set up from the FP32 values alone. Needs a plan whose every
 * dof is one of 3 consecutive dofs of a node (a row node).
 * EWFE_F32_MFMA_D64 (ewfe_eval_assemble only): the EWFE_F32_MFMA tangent stored as EWFE_F32_D64; every other call
 * takes such values as EWFE_F32_D64. No counterpart in the reference (its tangent is FP64). */
enum ewfe_dtype {
    EWFE_F64 = 0,
    EWFE_F32 = 1,
    EWFE_F32_MFMA = 2,
    EWFE_F16 = 3, /* ewfe_mg_set_fine_precision only */
    EWFE_F32_D64 = 4,
    EWFE_F32_MFMA_D64 = 5 /* ewfe_eval_assemble only */
};

/* Error codes */
#define EWFE_OK 0
#define EWFE_EINVAL -1   /* bad argument */
#define EWFE_EHIP -2     /* HIP runtime error */
#define EWFE_ENOMEM -3   /* allocation failed */
#define EWFE_ERANGE -4   /* size beyond int32 CSR indexing, or beyond a fixed scratch size */

/* Size of the fixed device scratch of ewfe_pcg_noalloc (64 MiB: systems of up to about 1.19 million dofs). */
#define EWFE_PCG_NOALLOC_MAX_BYTES (64ull << 20)

const char *ewfe_last_error(void);
int ewfe_version(void);

/* Builds the CSR pattern, the element -> CSR slot map and the element colouring on the host,
 * then uploads the plan's index structures to the current device.
 *   conn_host      int32 [n_el * npe]  node indices of each element, element-local node order
 *   node_dof_host  int64 [n_nodes]     first dof of each node (its 3 dofs are consecutive), -1 = none
 *   n_dof          size of the dof vector
 * Outputs nnz; the pattern itself is read with ewfe_plan_pattern. */
int ewfe_plan_create(ewfe_plan **plan, int elem_kind, int64_t n_el, const int32_t *conn_host, int64_t n_nodes,
                     const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out);

/* ewfe_plan_create with a chosen number of pipeline chunks for 8-node elements (0 = automatic, currently
 * one chunk). With more than one chunk the element pass of chunk c + 1 runs beside the CSR gather of chunk
 * c on a second stream; results are bitwise independent of the count. Same role as ewfe_plan_create
 * (dofmanager.py:522-557, csrgenerator.pyx:47-98); the count itself has no counterpart in the reference. */
int ewfe_plan_create_chunked(ewfe_plan **plan, int elem_kind, int64_t n_el, const int32_t *conn_host,
                             int64_t n_nodes, const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out,
                             int pipeline_chunks);

/* Assembly path of a 20-node plan (no counterpart in the reference; tools and tests): use_gather 1 = element records
 * summed by the row-node gather (the default where the plan has it: every node in at most 8 elements and coupled
 * with at most 128 nodes), 0 = colour-ordered additions into the CSR values. Both give the same sums up to the
 * order of the floating-point additions. EWFE_EINVAL for 1 on a plan without the gather lists; 8-node plans always
 * gather and take 1 only. */
int ewfe_plan_set_row_gather(ewfe_plan *plan, int32_t use_gather);

/* Pipeline chunks in use (0 for 20-node elements) and whether the mesh was recognised as a boxGen
 * structured hexahedral grid (boxgen.py:125-185), which selects the tile-streaming work orders. */
int ewfe_plan_pipeline_info(const ewfe_plan *plan, int32_t *n_chunks, int32_t *structured);

/* Same as ewfe_plan_create_chunked without touching any device: the plan answers ewfe_plan_pattern,
 * ewfe_plan_element_order, ewfe_plan_info, ewfe_plan_pipeline_info and ewfe_plan_pipeline_orders only
 * (host-side checks and tools). */
int ewfe_plan_create_host(ewfe_plan **plan, int elem_kind, int64_t n_el, const int32_t *conn_host, int64_t n_nodes,
                          const int64_t *node_dof_host, int64_t n_dof, int64_t *nnz_out, int pipeline_chunks);

/* Host copies of the 8-node pipeline orders (tools and tests; no counterpart in the reference):
 * pipe_el, pipe_row int64[n_chunks + 1] chunk boundaries; eorder int32[n_el] element indices in
 * evaluation order (= the plan order of ewfe_plan_element_order); rorder int32[n_row_nodes] row nodes
 * (ascending first dof = row node index) in gather order. Any pointer may be NULL. */
int ewfe_plan_pipeline_orders(const ewfe_plan *plan, int64_t *pipe_el_host, int64_t *pipe_row_host,
                              int32_t *eorder_host, int32_t *rorder_host);

/* Copies the CSR pattern to host buffers: indptr int32[n_dof+1], indices int32[nnz]. It equals
 * scipy's canonical COO->CSR of the VIJ pattern bit for bit (csrgenerator.pyx:68-77). */
int ewfe_plan_pattern(const ewfe_plan *plan, int32_t *indptr_host, int32_t *indices_host);

/* Device copies of the pattern (owned by the plan, valid until destroy). */
int ewfe_plan_pattern_dev(const ewfe_plan *plan, const int32_t **indptr_dev, const int32_t **indices_dev);

/* Bytes of a CSR value array of the plan in precision vals_dtype (8 nnz for EWFE_F64, 4 nnz for EWFE_F32 and
 * EWFE_F32_MFMA, the layout described at enum ewfe_dtype for EWFE_F32_D64); EWFE_EINVAL for another dtype or for
 * EWFE_F32_D64 on a plan without row nodes covering every dof. No counterpart in the reference. */
int64_t ewfe_vals_bytes(const ewfe_plan *plan, int vals_dtype);

/* Element order of the plan's quadrature-point state arrays: order_host[p] = element index stored at
 * position p (state layout is [component][qp][p], structure of arrays). n_colors_out may be NULL. */
int ewfe_plan_element_order(const ewfe_plan *plan, int64_t *order_host, int32_t *n_colors_out);

/* Number of state doubles per quadrature point the kernels keep for a material:
 * linear elastic 0, von Mises 7 (plastic strain 6 in engineering Voigt [11,22,33,12,13,23] + kappa),
 * Neo-Hooke 0 (path independent). */
int ewfe_state_size(int mat_kind);
int ewfe_plan_info(const ewfe_plan *plan, int64_t *n_el, int32_t *nodes_per_el, int32_t *qp_per_el,
                   int64_t *n_dof);
int ewfe_plan_destroy(ewfe_plan *plan);

/* System plan of a model with several element blocks (element types or material kinds): the reference keeps all
 * elements of a model in one VIJ pattern (DofManager over model.elements, dofmanager.py:522-557) and one
 * CSRGenerator (csrgenerator.pyx:47-98). Here every block has its own plan over the common dofs (node_dof and n_dof
 * of every plan the same); the system plan holds the row-wise union of their patterns (sorted unique columns: the
 * canonical CSR of the union of the blocks' VIJ entries) and its row-node graph, and takes ewfe_dirichlet_rows,
 * ewfe_spmv, ewfe_residual, ewfe_pcg(_noalloc), ewfe_pcg_workspace_size and ewfe_mg_create like an element plan;
 * ewfe_eval_assemble and the load / result calls stay with the block plans. plans[0 .. n_plans) stay owned by the
 * caller and may be destroyed before the system plan. */
int ewfe_plan_create_union(ewfe_plan **plan, const ewfe_plan *const *plans, int32_t n_plans, int64_t n_nodes,
                           const int64_t *node_dof_host, int64_t *nnz_out);

/* dst[map(k)] += src[k] over the pattern of block `block` of a system plan: the block's CSR values (its plan's
 * pattern) added into the system's CSR values. Blocks added in block order give every system entry the sum of its
 * block contributions in that order (csrgenerator.pyx:100-115 sums them in element order). */
int ewfe_csr_accumulate(const ewfe_plan *system_plan, int32_t block, const void *src_vals_dev, void *dst_vals_dev,
                        int vals_dtype, void *stream);

/* Per-element material parameter sets (several sections with materials of one kind, e.g. two linear
 * elastic materials on two element sets: every section assigns its material to the elements of its sets,
 * ew/sections/base/sectionbase.py:84-112, ew/generators/abqmodelconstructor.py:326-370). params_host holds
 * n_mat sets of 8 doubles (the layout of mat_params_host below); material_of_element_host[e] (original element
 * order) selects element e's set. Once set, every evaluation, quadrature-point result and adjoint call of the
 * plan takes each element's parameters from its set and ignores mat_params_host; el_param_dev still overrides
 * the first parameter. n_mat = 0 clears the table. Fails with EWFE_EINVAL on an index outside [0, n_mat). */
int ewfe_plan_set_materials(ewfe_plan *plan, int32_t n_mat, const double *params_host,
                            const int32_t *material_of_element_host);

/* One Newton iteration's element evaluation + assembly (the reference zeroes P, F and the CSR values
 * first, nonlinearimplicitstatic.py:419, csrgenerator.pyx:111; so does this call).
 *   mat_params_host  f64[8]  linear elastic [E, nu]; von Mises [E, nu, fy0, H, dfy, delta];
 *                            Neo-Hooke Wa [mu, K]; St. Venant-Kirchhoff [E, nu]
 *   el_param_dev     f64[n_el] or NULL: per-element override of mat_params[0] (original element order)
 *   X_dev            f64[n_dof] reference coordinates laid out like a dof vector (X[dof(node)+c])
 *   U_np_dev, dU_dev f64[n_dof] current total displacement and increment displacement
 *   state_n_dev      f64[nstate * nqp * n_el] converged state, plan order; NULL if nstate == 0
 *   state_t_dev      same layout, receives the trial state
 *   P_dev, F_dev     f64[n_dof] out: P = -f_int (element "P"), F = sum |Pe|
 *   csr_vals_dev     out: the assembled tangent, vals_dtype EWFE_F64 (f64 [nnz]), EWFE_F32 or EWFE_F32_MFMA
 *                    (f32 [nnz]), EWFE_F32_D64 or EWFE_F32_MFMA_D64 (ewfe_vals_bytes(plan, EWFE_F32_D64) bytes)
 *   pnewdt_dev       f64[1] out: 1e36, or < 1 if a quadrature point requests a cutback
 */
int ewfe_eval_assemble(const ewfe_plan *plan, int mat_kind, const double *mat_params_host,
                       const double *el_param_dev, const double *X_dev, const double *U_np_dev,
                       const double *dU_dev, const double *state_n_dev, double *state_t_dev, double *P_dev,
                       double *F_dev, void *csr_vals_dev, int vals_dtype, double *pnewdt_dev, void *stream);

/* Dirichlet rows of the assembled matrix: zero the rows listed in rows_dev (int32 [n_rows]) and put
 * 1 on their diagonal (nonlinearimplicitstatic.py:576-584; the reference's eliminate_zeros acts on a
 * copy and does not change the generator's pattern). */
int ewfe_dirichlet_rows(const ewfe_plan *plan, const int32_t *rows_dev, int64_t n_rows, void *csr_vals_dev,
                        int vals_dtype, void *stream);

/* Adjoint sensitivity of a per-element material parameter (no counterpart in the reference, whose solvers
 * have no adjoint; the per-element parameter mechanism is ew/sections/base/sectionbase.py:86-128):
 *   grad_dev[e] = lambda^T dR/dtheta_e,  R = P + PExt the Newton residual, theta_e = the element's first
 *   material property (el_param_dev[e] if given, else mat_params[0]): mu of Neo-Hooke Wa on C3D8TL, E of
 *   linear elasticity on C3D8. With K lambda = dJ/dU on the free dofs (lambda = 0 on Dirichlet dofs) this
 *   is dJ/dtheta_e of an objective J(U) at equilibrium. Element kinematics are recomputed. */
int ewfe_adjoint_param_grad(const ewfe_plan *plan, int mat_kind, const double *mat_params_host,
                            const double *el_param_dev, const double *X_dev, const double *U_dev,
                            const double *lambda_dev, double *grad_dev, void *stream);

/* y = K x over the plan's CSR pattern (K in f64 or f32, x and y f64). */
int ewfe_spmv(const ewfe_plan *plan, const void *csr_vals_dev, int vals_dtype, const double *x_dev, double *y_dev,
              void *stream);

/* r = b - K x (K in f64 or f32, b, x, r f64) with every row's sum accurate to twice the working precision
 * (compensated dot product). The residual of iterative refinement for NIST.linearSolve
 * (nonlinearimplicitstatic.py:727-751): the reference's direct solvers refine their solutions (PARDISO, the
 * default of edelweissfe/config/linsolve.py:46-86, runs iterative-refinement steps by default); with residuals in
 * twice the precision the refined solution is accurate to working precision however ill-conditioned K is. */
int ewfe_residual(const ewfe_plan *plan, const void *csr_vals_dev, int vals_dtype, const double *b_dev,
                  const double *x_dev, double *r_dev, void *stream);

/* Bytes of device workspace ewfe_pcg needs for this plan. */
int64_t ewfe_pcg_workspace_size(const ewfe_plan *plan);

/* Solves K x = b for the Newton correction, K being the assembled matrix after ewfe_dirichlet_rows with
 * the same rows (NIST.linearSolve, nonlinearimplicitstatic.py:727-751; the reference factorises the
 * row-eliminated matrix, this call runs Jacobi-preconditioned CG on its symmetric free-free block, which
 * has the same solution). Stops when ||b - K x||_2 <= rtol ||b||_2 or after maxit iterations; the host
 * polls for convergence every check_every iterations. Reports iterations and the final relative residual.
 * Fails with EWFE_EINVAL if p.Kp <= 0 (matrix not positive definite on the free dofs). */
int ewfe_pcg(const ewfe_plan *plan, const void *csr_vals_dev, int vals_dtype, const double *b_dev, double *x_dev,
             const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every, void *workspace_dev,
             int *iters_out, double *relres_out, void *stream);

/* ewfe_pcg without a workspace argument: the workspace is a fixed device array of EWFE_PCG_NOALLOC_MAX_BYTES inside
 * the library (one per GPU, held from the library's load to the end of the process). No counterpart in the
 * reference: this entry point is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); it calls no allocating routine. Same arguments, stopping
 * rule, results and errors as ewfe_pcg (bitwise the same iterates); fails with EWFE_ERANGE when
 * ewfe_pcg_workspace_size(plan) exceeds EWFE_PCG_NOALLOC_MAX_BYTES. Every call on one GPU shares its array:
 * calls that overlap on one GPU (several host threads or streams) race. */
int ewfe_pcg_noalloc(const ewfe_plan *plan, const void *csr_vals_dev, int vals_dtype, const double *b_dev,
                     double *x_dev, const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                     int *iters_out, double *relres_out, void *stream);

/* Smoothed-aggregation multigrid preconditioner for the same systems (the reference's iterative solver is
 * GMRES with a pyamg smoothed-aggregation preconditioner, edelweissfe/linsolve/gmres/gmres.py:70-88, selected
 * by name in edelweissfe/config/linsolve.py:46-86; here: a device-resident SA V-cycle inside CG).
 * ewfe_mg_create builds the aggregation and every sparsity pattern on the host (once per plan; needs 3 dofs
 * on every node) and allocates the hierarchy on the plan's device; degree = Chebyshev smoother degree,
 * ratio = lambda_max / lambda_min of the smoothing interval. The plan must outlive the hierarchy. */
int ewfe_mg_create(const ewfe_plan *plan, int degree, double ratio, ewfe_mg **mg_out, int *n_levels_out);
void ewfe_mg_destroy(ewfe_mg *mg);
int ewfe_mg_level_info(const ewfe_mg *mg, int level, int64_t *n_nodes, int64_t *nnz_blocks);

/* Numeric setup for one matrix (after ewfe_dirichlet_rows with the same rows): block-Jacobi inverses,
 * lambda_max(D^-1 A) per level by power iteration, prolongators, Galerkin products, dense coarsest inverse
 * (one host synchronisation). */
int ewfe_mg_setup(ewfe_mg *mg, const void *csr_vals_dev, int vals_dtype, const int32_t *rows_dev, int64_t n_rows,
                  void *stream);
/* lambda_max(D^-1 A) of every level, after ewfe_mg_setup (host array of n >= levels). */
int ewfe_mg_lambda(const ewfe_mg *mg, double *lambda_out, int n);
/* Precision of the V-cycle's fine-level passes (smoothing steps and residuals). EWFE_F64 / EWFE_F32 (the default):
 * the matrix values passed to each apply. EWFE_F16: every later ewfe_mg_setup also stores the set-up matrix divided
 * by a power of two (its largest magnitude then in [2^14, 2^15)) and rounded to FP16, and the V-cycle's fine-level
 * passes read that copy (half the bytes of FP32 values); the block-diagonal inverses, spectral bounds, prolongators
 * and coarse operators still come from the set-up matrix in its own precision, and the CG iteration around the
 * V-cycle keeps the caller's values. The copy (2 bytes per nonzero) is allocated here, not in the setup. No
 * counterpart in the reference (its preconditioner is pyamg's, in FP64). */
int ewfe_mg_set_fine_precision(ewfe_mg *mg, int dtype);

/* ewfe_pcg with the multigrid V-cycle as preconditioner; the hierarchy must be set up for csr_vals_dev.
 * Same arguments, stopping rule, workspace (ewfe_pcg_workspace_size) and errors as ewfe_pcg. */
int ewfe_pcg_mg(ewfe_mg *mg, const void *csr_vals_dev, int vals_dtype, const double *b_dev, double *x_dev,
                const int32_t *rows_dev, int64_t n_rows, double rtol, int maxit, int check_every,
                void *workspace_dev, int *iters_out, double *relres_out, void *stream);

/* Body force (edelweissfe/elements/displacementelement/element.py:348-371, assembled per element set by
 * edelweissfe/solvers/nonlinearimplicitstatic.py:903-945): PExt += sum over the elements with
 * el_mask_dev[element] != 0 (all elements when NULL; mask indexed by element) of
 * P_e,a = sum_q N_a(xi_q) b detJ_q w_q, b = force_host[0..2], reference configuration, summed per node in
 * ascending element order. Allocates its scratch in the plan on first use (8 npe bytes per element). */
int ewfe_body_force(ewfe_plan *plan, const double *X_dev, const uint8_t *el_mask_dev, const double *force_host,
                    double *PExt_dev, void *stream);

/* Quadrature-point results at displacement U_dev and state state_dev (plan order, as ewfe_eval_assemble;
 * J2 stress only): out_dev [n_el][qp][6] in element order; result 0 = "stress", 1 = "strain", the arrays the
 * reference's elements expose through getResultArray (displacementelement/element.py:222-235,
 * displacementtlelement/element.py:377-412) and ElementResultCollector gathers
 * (edelweissfe/utils/elementresultcollector.pyx:37-127). Small strain: Voigt [11,22,33,12,13,23], total
 * Lagrangian (Kirchhoff stress, Green strain): [11,22,33,12,23,13]; engineering shear strains. */
int ewfe_qp_results(const ewfe_plan *plan, int mat_kind, const double *mat_params_host, const double *el_param_dev,
                    const double *X_dev, const double *U_dev, const double *state_dev, int result, double *out_dev,
                    void *stream);

/* Distributed load of type pressure (edelweissfe/solvers/nonlinearimplicitstatic.py:460-501 with the
 * stepaction edelweissfe/stepactions/distributedload.py; the element side is the Marmot element's
 * computeDistributedLoad, element.pxd): PExt += -p * integral over face `face` (Abaqus numbering 1..6) of
 * N_a n dA, reference configuration, for the elements elements_dev[0..n) (element indices; an element listed
 * twice is loaded once, indices outside [0, n_el) are skipped); summed per node in ascending element order.
 * Allocates its scratch in the plan on first use (8 bytes per element, 24 npe bytes per listed element). */
int ewfe_surface_pressure(ewfe_plan *plan, const double *X_dev, const int32_t *elements_dev, int64_t n_elements,
                          int face, double pressure, double *PExt_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
