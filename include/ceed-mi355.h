/*
 * ceed-mi355.h -- host-side entry points of the /gpu/hip/mi355 backend.
 *
 * The fused kernels are reached through the thin C ABI in
 * ceed-mi355-kernels.h; this header holds the host helpers the backend and
 * its users need (quadrature, 1D Lagrange matrices).
 */
#ifndef CEED_MI355_H
#define CEED_MI355_H

#include <ceed/ceed.h>

#include "ceed-mi355-kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre points/weights on [-1, 1]; same signature and result as the
 * reference's public CeedGaussQuadrature (interface/ceed-basis.c:2529). */
int CeedGaussQuadrature(CeedInt Q, CeedScalar *q_ref_1d, CeedScalar *q_weight_1d);
/* Gauss-Lobatto points/weights; q_weight_1d may be NULL.  Replaces
 * CeedLobattoQuadrature (interface/ceed-basis.c:2581). */
int CeedLobattoQuadrature(CeedInt Q, CeedScalar *q_ref_1d, CeedScalar *q_weight_1d);
/* 1D Lagrange matrices on P Gauss-Lobatto nodes evaluated at Q points
 * (quad_mode 0 = Gauss, 1 = Gauss-Lobatto): interp_1d, grad_1d are Q x P row
 * major.  The matrices CeedBasisCreateTensorH1Lagrange builds
 * (interface/ceed-basis.c:1617-1690). */
int CeedMi355LagrangeBasis1D(int32_t P, int32_t Q, int32_t quad_mode, double *interp_1d, double *grad_1d, double *q_ref_1d, double *q_weight_1d);

/* Backend registration: adds "/gpu/hip/mi355" to the resource registry, as
 * the reference's backends/hip-ref/ceed-hip-ref.c:75 registers "/gpu/hip/ref"
 * (CEED_INTERN int CeedRegister_Hip(void)).  Called by the library's
 * registry set-up; a reference build would call it from
 * backends/ceed-backend-list.h. */
int CeedRegister_Mi355(void);
/* Backend init for a resource "/gpu/hip/mi355[:device_id=N]"; the
 * counterpart of CeedInit_Hip_ref (backends/hip-ref/ceed-hip-ref.c:43-73). */
int CeedInit_Mi355(const char *resource, Ceed ceed);

#ifdef __cplusplus
}
#endif

#endif /* CEED_MI355_H */
