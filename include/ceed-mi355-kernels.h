/*
 * ceed-mi355-kernels.h -- thin C ABI from host code to the hand-written HIP
 * kernels for gfx950 (MI355X).  Plain pointers and sizes only.
 *
 * Every entry point is an MI355X-native replacement for one step of the
 * reference hot path v = E^T B^T D B E u; the comment on each names the
 * reference function (file:line in CEED/libCEED) whose result it reproduces.
 *
 * Conventions
 *   - Return 0 on success, a negative CEED_MI355_ERROR_* code on failure; the
 *     message is available from CeedMi355GetLastError() on the same thread.
 *   - Pointers named *_host are host memory; every other array argument is
 *     device memory (hipMalloc'ed or a torch CUDA tensor's data_ptr()).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - Modes use the reference's enum values: t_mode 0 = NOTRANSPOSE,
 *     1 = TRANSPOSE; eval_mode 1 = INTERP, 2 = GRAD, 16 = WEIGHT
 *     (include/ceed/types.h:196-247).
 *   - E-vectors are element-major [elem][comp][node]; basis Q-vectors are
 *     [elem][dim][comp][qpt] (GRAD) or [elem][comp][qpt] (INTERP).
 */
#ifndef CEED_MI355_KERNELS_H
#define CEED_MI355_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes share the values of CeedErrorType (include/ceed/types.h:162-181). */
enum {
  CEED_MI355_SUCCESS           = 0,
  CEED_MI355_ERROR_DIMENSION   = 2,
  CEED_MI355_ERROR_MAJOR       = -1,
  CEED_MI355_ERROR_BACKEND     = -2,
  CEED_MI355_ERROR_UNSUPPORTED = -3,
};

/* Arithmetic type of an operator plan. */
enum { CEED_MI355_FP64 = 0, CEED_MI355_FP32 = 1 };

/* Pointwise operators the fused kernel inlines (the gallery QFunctions
 * MassApply / Vector3MassApply, Poisson3DApply / Vector3Poisson3DApply, and a
 * kappa-weighted Poisson3DApply for -div(kappa grad u)). */
enum { CEED_MI355_QF_MASS = 0, CEED_MI355_QF_DIFFUSION = 1, CEED_MI355_QF_DIFFUSION_KAPPA = 2 };
/* Field kinds of a run-time compiled fused operator (CeedMi355OperatorCreateHexJit):
 * an active field evaluated with CEED_EVAL_INTERP or CEED_EVAL_GRAD, or the
 * passive CEED_EVAL_NONE input (the plan's q_data). */
enum { CEED_MI355_FIELD_INTERP = 0, CEED_MI355_FIELD_GRAD = 1, CEED_MI355_FIELD_PASSIVE = 2 };

/* E^T mode flags of CeedMi355OperatorCreateHex (0: transpose map for P >= 6, colours below). */
enum {
  CEED_MI355_PLAN_COLOR   = 2,
  CEED_MI355_PLAN_ATOMIC  = 4,
  CEED_MI355_PLAN_TMAP    = 8,
  CEED_MI355_PLAN_SPLIT   = 16,
  CEED_MI355_PLAN_BRICK   = 32,
  CEED_MI355_PLAN_SEGMENT = 64
};

/* Geometric-factor builders (gallery Mass3DBuild / Poisson3DBuild). */
enum { CEED_MI355_BUILD_MASS = 0, CEED_MI355_BUILD_DIFFUSION = 1 };

/* ------------------------------------------------------------------------ */
/* Device and memory                                                         */
/* ------------------------------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char *CeedMi355GetLastError(void);
/* Select the device and check it is a gfx950; replaces the device set-up of
 * backends/hip/ceed-hip-common.c:18-40 (CeedInit_Hip). */
int CeedMi355DeviceInit(int device_id);
int CeedMi355Malloc(size_t bytes, void **ptr);
int CeedMi355Free(void *ptr);
int CeedMi355Memcpy(void *dst, const void *src, size_t bytes, void *stream); /* any direction, stream-ordered */
int CeedMi355Memset(void *dst, int value, size_t bytes, void *stream);
int CeedMi355StreamSynchronize(void *stream);

/* ------------------------------------------------------------------------ */
/* Element restriction (E / E^T)                                             */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355Restriction_private *CeedMi355Restriction;

/* Offset restriction.  Copies the offsets to the device and builds the
 * transpose map used by the bit-exact E^T.  Replaces CeedElemRestrictionCreate
 * on /cpu/self/ref/serial (backends/ref/ceed-ref-restriction.c:746-905). */
int CeedMi355RestrictionCreateOffset(int32_t num_elem, int32_t elem_size, int32_t num_comp, int32_t comp_stride, int64_t l_size,
                                     const int32_t *offsets_host, CeedMi355Restriction *rstr);
/* Strided restriction: L index = node*strides[0] + comp*strides[1] + elem*strides[2]. */
int CeedMi355RestrictionCreateStrided(int32_t num_elem, int32_t elem_size, int32_t num_comp, int64_t l_size, const int32_t *strides_host,
                                      CeedMi355Restriction *rstr);
/* NOTRANSPOSE: v(E-vector) = E u, replaces ceed-ref-restriction.c:19-75.
 * TRANSPOSE:   v(L-vector) += E^T u, summed per L entry in ascending
 *              (elem, comp, node) order, bit-identical to
 *              ceed-ref-restriction.c:181-242 (serial, block size 1). */
int CeedMi355RestrictionApply(CeedMi355Restriction rstr, int t_mode, const double *u, double *v, void *stream);
int CeedMi355RestrictionDestroy(CeedMi355Restriction *rstr);

/* ------------------------------------------------------------------------ */
/* Tensor-product H1 basis (B / B^T), dim 1..3                                */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355Basis_private *CeedMi355Basis;

/* interp_1d and grad_1d are Q x P row-major host arrays; q_weight_1d has Q
 * entries.  Replaces CeedBasisCreateTensorH1_Ref (ceed-ref-basis.c:283-310). */
int CeedMi355BasisCreateTensorH1(int32_t dim, int32_t num_comp, int32_t P_1d, int32_t Q_1d, const double *interp_1d_host,
                                 const double *grad_1d_host, const double *q_weight_1d_host, CeedMi355Basis *basis);
/* Replaces CeedBasisApply / CeedBasisApplyAdd (ceed-ref-basis.c:19-257;
 * interface/ceed-basis.c:2083).  add != 0 sums into v. */
int CeedMi355BasisApply(CeedMi355Basis basis, int32_t num_elem, int t_mode, int eval_mode, int add, const double *u, double *v,
                        void *stream);
int CeedMi355BasisDestroy(CeedMi355Basis *basis);

/* ------------------------------------------------------------------------ */
/* Fused hex operator  v (+)= E^T B^T D B E u                                 */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355Operator_private *CeedMi355Operator;

/* Plan for a 3D tensor hex operator with a gallery pointwise D, applied by
 * one fused kernel launch.  E^T (flags; 0 picks the mode measured fastest on
 * MI355X for the order, precision and size: COLOR for P_1d <= 5 from 10M
 * E-vector entries, SPLIT for P_1d = 7, 8 in FP64 with one component from
 * 10M entries, TMAP otherwise; CEED_MI355_SCATTER=color|tmap|split|atomic|brick|segment
 * replaces the choice):
 *   CEED_MI355_PLAN_TMAP   the fused kernel stores every contribution into
 *                          an E-vector sorted by L entry (slots from the
 *                          transpose offset map); a gather kernel sums each
 *                          L entry's contiguous run in ascending
 *                          (elem, comp, node) order: no atomics, bitwise
 *                          reproducible, the reference's summation order
 *                          (ceed-ref-restriction.c:220-242);
 *   CEED_MI355_PLAN_COLOR  elements coloured so that no two of a colour share
 *                          an L entry; one launch per colour adds in place
 *                          (no atomics, bitwise reproducible);
 *   CEED_MI355_PLAN_SPLIT  entries reached by at most two contributions are
 *                          summed by atomic adds onto a zeroed target (a
 *                          two-term sum does not depend on the order), the
 *                          others by the transpose map restricted to them:
 *                          every entry summed in the reference's order, Apply
 *                          bitwise equal to TMAP; ApplyAdd forms A u in a
 *                          scratch vector, then adds it;
 *   CEED_MI355_PLAN_ATOMIC hardware atomic adds, sums in arrival order;
 *   CEED_MI355_PLAN_BRICK  a workgroup per brick of up to 2x2x2 face-sharing
 *                          elements, the brick's shared nodes summed in LDS,
 *                          brick-surface nodes finished by an ordered gather
 *                          (mass and diffusion only);
 *   CEED_MI355_PLAN_SEGMENT a workgroup per segment of up to 256 / Q_1d^2
 *                          (or CEED_MI355_SEGMENT_MAX, if smaller)
 *                          elements chained along x through shared faces;
 *                          the face between consecutive elements is summed in
 *                          LDS (left element's value first), segments are
 *                          coloured so that no two of a colour share an L
 *                          entry, one launch per colour adds in place (no
 *                          atomics, bitwise reproducible; mass and diffusion
 *                          Apply / ApplyAdd only, line thread layout).
 * Supports Q_1d in {P_1d, P_1d + 1}, P_1d in 2..9.  scalar_type selects FP64
 * or FP32 storage and arithmetic for u, v and q_data.  Replaces the
 * per-operator set-up of ceed-ref-operator.c:145-210. */
int CeedMi355OperatorCreateHex(int qf, int scalar_type, int32_t num_comp, int32_t num_elem, int32_t P_1d, int32_t Q_1d,
                               const double *interp_1d_host, const double *grad_1d_host, const int32_t *offsets_host, int32_t comp_stride,
                               int64_t l_size, int flags, CeedMi355Operator *op);
/* A fused plan like CeedMi355OperatorCreateHex whose pointwise operator is a
 * user QFunction: the source file (CEED_QFUNCTION(function_name), local
 * includes inlined) is compiled with hipRTC into the line-layout fused kernel,
 * which hands the QFunction its fields in their order (in_kinds / out_kinds:
 * CEED_MI355_FIELD_*) one point at a time.  Scalar active fields (INTERP
 * and/or GRAD, in and out), at most one passive input of q_comp values per
 * point (set with CeedMi355OperatorSetQData), FP64.  The counterpart of the
 * reference's hip-gen operator, which generates a fused kernel around any
 * QFunction (backends/hip-gen/ceed-hip-gen-operator-build.cpp). */
int CeedMi355OperatorCreateHexJit(const char *source_path, const char *function_name, int32_t num_in, const int32_t *in_kinds_host,
                                  int32_t num_out, const int32_t *out_kinds_host, int32_t q_comp, int32_t num_elem, int32_t P_1d, int32_t Q_1d,
                                  const double *interp_1d_host, const double *grad_1d_host, const int32_t *offsets_host, int64_t l_size, int flags,
                                  CeedMi355Operator *op);
/* Device pointer of the QFunction context for later applies of a plan made by
 * CeedMi355OperatorCreateHexJit (NULL: none). */
int CeedMi355OperatorSetContext(CeedMi355Operator op, void *ctx_device);
/* Geometric factors from mesh coordinates on the GPU, FP64 arithmetic:
 * Mass3DBuild (w det J) or Poisson3DBuild (6 Voigt entries of
 * w det J J^-1 J^-T).  The mesh restriction has mesh_P^3 nodes per element,
 * 3 components at mesh_comp_stride; coords is an FP64 L-vector.  Replaces
 * applying the build operator of examples/ceed/ex1-volume.c:195-236 with the
 * gallery QFunctions jit-source/gallery/ceed-{mass3d,poisson3d}build.h. */
int CeedMi355OperatorBuildQData(CeedMi355Operator op, int build, int32_t mesh_P_1d, const double *mesh_interp_1d_host,
                                const double *mesh_grad_1d_host, const double *q_weight_1d_host, const int32_t *mesh_offsets_host,
                                int32_t mesh_comp_stride, const double *coords, void *stream);
/* Load geometric factors from a caller array q_data (op scalar type) in the
 * backend-strided layout [elem][comp][qpt] of CEED_STRIDES_BACKEND. */
int CeedMi355OperatorSetQData(CeedMi355Operator op, const void *q_data, void *stream);
/* Copy the geometric factors back out in [elem][comp][qpt] layout (FP64 or FP32 per the plan). */
int CeedMi355OperatorGetQData(CeedMi355Operator op, void *q_data, void *stream);
/* kappa at quadrature points, [elem][qpt] FP64 (QF_DIFFUSION_KAPPA only). */
int CeedMi355OperatorSetKappa(CeedMi355Operator op, const double *kappa, void *stream);
/* Adjoint of a QF_DIFFUSION_KAPPA plan (FP64, scalar), one fused kernel:
 *   v = A(kappa)^T lambda                       (overwritten; A is symmetric)
 *   dkappa[e][q] = grad(lambda)^T S grad(u)     at every quadrature point,
 * S = w det J J^-1 J^-T, the derivative of lambda^T A(kappa) u in kappa_q;
 * dkappa uses the [elem][Q^3] layout of CeedMi355OperatorSetKappa.  For a
 * loss L(A(kappa) u), lambda = dL/d(Au) gives dL/du = v and dL/dkappa = dkappa.
 * The transpose apply is the reference's CeedOperatorApply of the same
 * operator (A symmetric; ceed-ref-operator.c:381-470). */
int CeedMi355OperatorApplyAdjoint(CeedMi355Operator op, const double *u, const double *lambda, double *v, double *dkappa, void *stream);
/* Diagonal of the plan's operator into the L-vector diag (plan scalar type);
 * add != 0 sums into diag.  Replaces CeedOperatorLinearAssembleDiagonal /
 * AddDiagonal (interface/ceed-preconditioning.c:2395-2480) for this operator. */
int CeedMi355OperatorAssembleDiagonal(CeedMi355Operator op, void *diag, int add, void *stream);
/* v += A u  (CeedOperatorApplyAdd, interface/ceed-operator.c:2313-2340). */
int CeedMi355OperatorApplyAdd(CeedMi355Operator op, const void *u, void *v, void *stream);
/* v  = A u  (CeedOperatorApply, interface/ceed-operator.c:2271-2294). */
int CeedMi355OperatorApply(CeedMi355Operator op, const void *u, void *v, void *stream);
/* Elements per workgroup of the fused kernel (1 <= epb, epb * Q^2 <= 256
 * threads, LDS <= 160 KiB, the gfx950 workgroup limit). */
int CeedMi355OperatorSetLaunch(CeedMi355Operator op, int32_t elems_per_block);
/* Thread layout of the fused kernel: 0 = z-columns per thread, x and y
 * contractions read point by point from LDS; 1 = whole 1D lines per thread
 * for the x and y contractions (about a third of the LDS traffic); 2 = a
 * thread per quadrature point, all three contractions through LDS (Q^3 <=
 * 256; not for run-time compiled QFunctions).  All give identical results.
 * elems_per_block <= 0 picks the default.  No
 * counterpart in the reference: its hip-shared/hip-gen backends fix the
 * thread layout per (P, Q) (backends/hip-shared/ceed-hip-shared-basis.c). */
int CeedMi355OperatorSetLayout(CeedMi355Operator op, int32_t layout, int32_t elems_per_block);
/* Current thread layout and elements per workgroup (either pointer may be NULL). */
int CeedMi355OperatorGetLaunch(CeedMi355Operator op, int32_t *layout, int32_t *elems_per_block);
/* Autotuner (the reference has none: its hip-shared/hip-gen backends fix the
 * launch shape per P, Q): time Apply (reps times) for every admissible
 * (thread layout, elements per workgroup) on scratch vectors, time the three
 * fastest again (twice each, 2 * reps applies, best time kept), keep the
 * fastest in the plan and remember it for later plans of the same (P, Q,
 * num_comp, pointwise operator, scalar type, E^T mode).  best_epb_host gets
 * the elements per workgroup; CeedMi355OperatorGetLaunch gives the layout. */
int CeedMi355OperatorAutotune(CeedMi355Operator op, int32_t reps, int32_t *best_epb_host, double *best_ms_host, void *stream);
/* Number of element colours, and the element count of each (host array of at
 * least num_colors entries, may be NULL). */
int CeedMi355OperatorGetColors(CeedMi355Operator op, int32_t *num_colors, int32_t *color_sizes_host);
/* The plan's E^T mode (CEED_MI355_PLAN_TMAP, _COLOR, _SPLIT or _ATOMIC), as
 * chosen at creation.  No counterpart in the reference (its backends have one
 * fixed E^T each). */
int CeedMi355OperatorGetScatter(CeedMi355Operator op, int32_t *mode);
int CeedMi355OperatorDestroy(CeedMi355Operator *op);

/* ------------------------------------------------------------------------ */
/* L-vector arithmetic                                                       */
/* ------------------------------------------------------------------------ */
enum {
  CEED_MI355_VEC_SET        = 0, /* z = alpha                       CeedVectorSetValue    (interface/ceed-vector.c:377) */
  CEED_MI355_VEC_SCALE      = 1, /* z = alpha z                     CeedVectorScale       (interface/ceed-vector.c:735) */
  CEED_MI355_VEC_AXPY       = 2, /* z = alpha x + z                 CeedVectorAXPY        (interface/ceed-vector.c:810) */
  CEED_MI355_VEC_AXPBY      = 3, /* z = alpha x + beta z            CeedVectorAXPBY       (interface/ceed-vector.c:881) */
  CEED_MI355_VEC_PMULT      = 4, /* z = x .* y                      CeedVectorPointwiseMult (interface/ceed-vector.c:953) */
  CEED_MI355_VEC_RECIPROCAL = 5, /* z = 1/z where |z| > alpha       CeedVectorReciprocal  (interface/ceed-vector.c:1042) */
};
enum {
  CEED_MI355_RED_DOT     = 0, /* sum x_i y_i  (x == y gives the squared 2-norm) */
  CEED_MI355_RED_ABS_SUM = 1, /* sum |x_i|    (CEED_NORM_1) */
  CEED_MI355_RED_MAX_ABS = 2, /* max |x_i|    (CEED_NORM_MAX) */
};
/* Elementwise vector operation in FP64 or FP32 (scalar_type).  Replaces the
 * /cpu/self/ref/serial vector kernels (backends/ref/ceed-ref-vector.c). */
int CeedMi355VectorOp(int op, int scalar_type, int64_t n, double alpha, double beta, const void *x, const void *y, void *z, void *stream);
/* Deterministic reduction into one FP64 device scalar: a fixed grid and a
 * fixed-order tree, so equal inputs give equal bits.  Replaces CeedVectorNorm
 * (interface/ceed-vector.c:677-733) and the dot products of the CG drivers. */
int CeedMi355VectorReduce(int op, int scalar_type, int64_t n, const void *x, const void *y, double *result_device, void *stream);
/* CeedMi355VectorReduce without any allocation, the same kernels and bits.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The block partials live in a fixed array of 1024
 * doubles (8 KiB) in each GPU's memory, held until the process ends; every n
 * fits it (the reduction grid is capped at 1024 blocks).  All calls on one GPU
 * share that array, so calls that overlap on one GPU race.  Nothing in the
 * array needs clearing: each call writes every slot it reads. */
int CeedMi355VectorReduceNoAlloc(int op, int scalar_type, int64_t n, const void *x, const void *y, double *result_device, void *stream);
/* y = (to_type) x, element by element (FP64 <-> FP32). */
int CeedMi355VectorConvert(int64_t n, int from_type, const void *x, int to_type, void *y, void *stream);

/* ------------------------------------------------------------------------ */
/* Run-time compiled QFunctions (hipRTC)                                     */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355QFunction_private *CeedMi355QFunction;

/* Compile the CEED_QFUNCTION `function_name` in the source file source_path
 * (local #include "..." files are inlined) into a pointwise kernel for
 * gfx950.  Replaces the QFunction JIT of backends/hip-ref/ceed-hip-ref-qfunction-load.cpp:23
 * (CeedQFunctionBuildKernel_Hip_ref). */
int CeedMi355QFunctionCompile(const char *source_path_host, const char *function_name_host, int32_t num_in, const int32_t *in_sizes_host,
                              int32_t num_out, const int32_t *out_sizes_host, CeedMi355QFunction *qf);
/* Evaluate at num_elem*Q points; field i, component k of point q of element e
 * is at in[i][(e*size_i + k)*Q + q] (in and out are host arrays of device
 * pointers).  ctx_device is the device copy of the QFunction context or NULL.
 * Replaces CeedQFunctionApply (interface/ceed-qfunction.c:1117). */
int CeedMi355QFunctionApply(CeedMi355QFunction qf, int64_t num_elem, int32_t Q, void *ctx_device, const double *const *in_host,
                            double *const *out_host, void *stream);
int CeedMi355QFunctionDestroy(CeedMi355QFunction *qf);

/* ------------------------------------------------------------------------ */
/* Conjugate gradients on a fused plan                                       */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355CG_private *CeedMi355CG;
typedef struct {
  int32_t iterations;             /* inner iterations run */
  int32_t replacements;           /* FP64 residual replacements (mixed precision) */
  double  rel_residual;           /* ||b - A x|| / ||b||, recomputed in FP64 at the end */
  double  rel_residual_recursive; /* the value the stopping test last saw */
} CeedMi355CGStats;

/* Unpreconditioned CG with device-resident scalars.  op64 == NULL: plain FP64
 * CG on the FP64 plan `op`.  op64 != NULL: `op` is an FP32 plan of the same
 * operator; the iteration runs in FP32 and every check_every iterations the
 * FP32 increment is added to the FP64 solution and the residual is replaced
 * by b - A64 x computed with op64.  The solver the reference's benchmark
 * problems run around CeedOperatorApply (PETSc KSPCG, examples/petsc/bps.c:210-239). */
int CeedMi355CGCreate(CeedMi355Operator op, CeedMi355Operator op64, CeedMi355CG *cg);
/* Solve A x = b from the given x (FP64 L-vectors).  Stops when
 * ||r|| <= rtol ||b|| (tested every check_every iterations; the recursive
 * residual for plain CG, the replaced FP64 residual for mixed precision) or
 * after max_iter iterations.  use_graph != 0 captures check_every iterations
 * in a hipGraph once and replays it.  Ordered after the caller's stream;
 * returns when the solve is done. */
int CeedMi355CGSolve(CeedMi355CG cg, const double *b, double *x, int32_t max_iter, double rtol, int32_t check_every, int use_graph, void *stream,
                     CeedMi355CGStats *stats_host);
/* Jacobi preconditioning for later solves: z = D^-1 r with D the operator's
 * diagonal (FP64, length l_size, e.g. from CeedMi355OperatorAssembleDiagonal;
 * entries must be non-zero); NULL switches it off.  The reference's PETSc
 * examples build this preconditioner from CeedOperatorLinearAssembleDiagonal
 * (examples/petsc/src/matops.c:54, interface/ceed-preconditioning.c:2395).
 * The stopping test stays on the unpreconditioned residual. */
int CeedMi355CGSetJacobi(CeedMi355CG cg, const double *diag, void *stream);
int CeedMi355CGDestroy(CeedMi355CG *cg);

/* ------------------------------------------------------------------------ */
/* Batched CG over independent systems                                       */
/* ------------------------------------------------------------------------ */
typedef struct CeedMi355BatchCG_private *CeedMi355BatchCG;

/* num_systems independent SPD systems on one concatenated vector: system s
 * owns entries [sys_begin[s], sys_begin[s+1]) (sys_begin has num_systems + 1
 * entries, sys_begin[0] = 0).  The block-diagonal operator is the list of FP64
 * plans ops[i], plan i acting on entries [op_begin[i], op_begin[i] +
 * l_size_i); the plan ranges tile the vector, and a plan may hold the
 * elements of several systems (one plan per (operator, P, Q) group).  No
 * counterpart in the reference, whose operators are applied one at a time;
 * this batches the CG solves of the benchmark problems
 * (examples/petsc/bps.c:210-239) over many meshes. */
int CeedMi355BatchCGCreate(int32_t num_ops, const CeedMi355Operator *ops_host, const int64_t *op_begin_host, int32_t num_systems,
                           const int64_t *sys_begin_host, CeedMi355BatchCG *cg);
/* Solve every system from the given x until ||r_s|| <= rtol ||b_s|| (per
 * system; a converged system is frozen) or max_iter iterations.  Checks every
 * check_every iterations; use_graph captures check_every iterations in a
 * hipGraph.  Per system: iterations run, the true relative residual, the
 * index of the check (0 = before the first iteration) at which it was first
 * seen converged (-1 if never), and (converged_ms_host, may be NULL) the host
 * time of that check in ms from the call's start (NaN if never). */
int CeedMi355BatchCGSolve(CeedMi355BatchCG cg, const double *b, double *x, int32_t max_iter, double rtol, int32_t check_every, int use_graph,
                          void *stream, int32_t *iterations_host, double *rel_residual_host, int32_t *converged_at_check_host,
                          int32_t *num_checks_host, double *converged_ms_host);
int CeedMi355BatchCGDestroy(CeedMi355BatchCG *cg);

typedef struct CeedMi355BatchSmall_private *CeedMi355BatchSmall;
/* Systems small enough for one workgroup, solved in one persistent launch:
 * each workgroup takes the next system off a work list (largest first) and
 * runs its whole unpreconditioned CG inside the workgroup, then takes the
 * next; no lockstep across systems, so each system's latency is its own
 * solve.  Same layout as CeedMi355BatchCGCreate, with every plan a scalar
 * FP64 colour plan (CEED_MI355_PLAN_COLOR) of mass or diffusion, Q = P + 1,
 * P = 2..9.  No counterpart in the reference (see CeedMi355BatchCGCreate).
 * Currently switched off: returns CEED_MI355_ERROR_UNSUPPORTED (its first GPU
 * test run faulted the GPU; the cause has not been found). */
int CeedMi355BatchSmallCreate(int32_t num_ops, const CeedMi355Operator *ops_host, const int64_t *op_begin_host, int32_t num_systems,
                              const int64_t *sys_begin_host, CeedMi355BatchSmall *bs);
/* x_s = A_s^-1 b_s for every system from x = 0, until ||r_s|| <= rtol ||b_s||
 * or max_iter iterations.  Per system (each may be NULL): iterations run, the
 * true relative residual, and the time in ms from the launch's start to the
 * system's completion (device clock).  Returns when the launch has ended. */
int CeedMi355BatchSmallSolve(CeedMi355BatchSmall bs, const double *b, double *x, int32_t max_iter, double rtol, void *stream,
                             int32_t *iterations_host, double *rel_residual_host, double *done_ms_host);
int CeedMi355BatchSmallDestroy(CeedMi355BatchSmall *bs);

#ifdef __cplusplus
}
#endif

#endif /* CEED_MI355_KERNELS_H */
