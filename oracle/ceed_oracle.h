/*
 * ceed_oracle.h -- CPU oracle for the tensor-product hex operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU backend /cpu/self/ref/serial (libCEED) for the hot path
 * v = E^T B^T D B E u.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product path
 * (libceed_amd) never calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Enum values equal the reference's (include/ceed/types.h:196-283).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   quadrature, Lagrange B/G    -- golden tests/output/t300-basis.out,
 *                                  closed forms, oracle/_ref bit comparison
 *   pseudoinverse/collocated    -- D_Q * B = G identity, oracle/_ref bits
 *   tensor contraction, basis   -- polynomial exactness, oracle/_ref bits
 *   restriction                 -- t200/t209 known answers, brute force
 *   gallery QFunctions          -- closed forms on affine maps, oracle/_ref
 *   operator                    -- volume/zero-row-sum/surface invariants,
 *                                  symmetry, oracle/_ref bits
 */
#ifndef CEED_ORACLE_H
#define CEED_ORACLE_H

#include <stdint.h>

enum { ORC_NOTRANSPOSE = 0, ORC_TRANSPOSE = 1 };
enum { ORC_EVAL_NONE = 0, ORC_EVAL_INTERP = 1, ORC_EVAL_GRAD = 2, ORC_EVAL_WEIGHT = 16 };
enum { ORC_GAUSS = 0, ORC_GAUSS_LOBATTO = 1 };

/* ---- 1D quadrature and Lagrange basis ---------------------------------- */
int orc_gauss_quadrature(int Q, double *q_ref, double *q_weight);
int orc_lobatto_quadrature(int Q, double *q_ref, double *q_weight);
int orc_lagrange_1d(int P, int Q, int quad_mode, double *interp, double *grad, double *q_ref, double *q_weight);
int orc_pseudoinverse(const double *mat, int m, int n, double *pinv);
int orc_collocated_grad(int P, int Q, const double *interp, const double *grad, double *collo);
int orc_is_collocated(int P, int Q, const double *interp);

/* ---- tensor contraction and basis apply -------------------------------- */
void orc_tensor_contract(int A, int B, int C, int J, const double *t, int t_mode, int add, const double *u, double *v);
int  orc_basis_apply(int dim, int num_comp, int P, int Q, const double *interp, const double *grad, const double *q_weight,
                     int num_elem, int t_mode, int eval_mode, int apply_add, const double *u, double *v);

/* ---- element restriction ------------------------------------------------ */
void orc_restriction_offset(int num_elem, int elem_size, int num_comp, int comp_stride, const int32_t *offsets, int t_mode,
                            const double *u, double *v);
void orc_restriction_strided(int num_elem, int elem_size, int num_comp, const int32_t *strides, int t_mode, const double *u,
                             double *v);

/* ---- gallery QFunctions (signature of CEED_QFUNCTION) ------------------ */
typedef int (*orc_qfunction)(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_mass3d_build(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_mass_apply(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_poisson3d_build(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_poisson3d_apply(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_vector3_mass_apply(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_vector3_poisson3d_apply(void *ctx, int Q, const double *const *in, double *const *out);
int orc_qf_diffusion_kappa_apply(void *ctx, int Q, const double *const *in, double *const *out);

/* ---- generic element-loop operator (ref/serial semantics) ------------- */
typedef struct {
  int eval_mode;   /* ORC_EVAL_* */
  int size;        /* QFunction field size (num_comp, times dim for GRAD) */
  /* restriction: strided when offsets == NULL */
  int            num_comp, elem_size, comp_stride;
  const int32_t *offsets;
  int32_t        strides[3];
  /* tensor H1 basis (INTERP, GRAD, WEIGHT) */
  int           dim, P, Q;
  const double *interp, *grad, *q_weight;
  /* L-vector: read for inputs, summed into for outputs */
  double *vec;
} orc_field;

int orc_operator_apply_add(int num_elem, int num_qpts, int num_in, const orc_field *in, int num_out, const orc_field *out,
                           orc_qfunction qf, void *ctx);

/* Flat wrappers for the hex configurations (callable through ctypes). */
/* kind: 0 = mass, 1 = diffusion */
int orc_hex_build_qdata(int kind, int num_elem, int P_mesh, int Q, int quad_mode, const int32_t *mesh_offsets, int mesh_comp_stride,
                        const double *coords, double *q_data);
/* num_comp: 1 or 3; the L-vector uses comp_stride for its components. */
int orc_hex_apply(int kind, int num_comp, int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, int comp_stride,
                  const double *q_data, const double *u, double *v);
/* Variable-coefficient diffusion: q_data holds 6 Voigt entries, kappa one value per quadrature point. */
int orc_hex_apply_kappa(int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, const double *q_data, const double *kappa,
                        const double *u, double *v);
/* Gradient of the kappa-weighted bilinear form: dk[e,q] = -sum_ij (grad u)_i S_ij (grad lam)_j. */
int orc_hex_kappa_gradient(int num_elem, int P, int Q, int quad_mode, const int32_t *offsets, const double *q_data, const double *u,
                           const double *lam, double *dkappa);

#endif
