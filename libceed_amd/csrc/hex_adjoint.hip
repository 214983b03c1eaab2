// Adjoint of the kappa-weighted diffusion operator A(kappa) u = E^T G^T (kappa S) G E u
// (S = w det J J^-1 J^-T, the Poisson3DBuild geometric factors), in one fused
// kernel per element:
//   v       = A(kappa)^T lambda = E^T B^T G^T (kappa S)^T G B E lambda  (S symmetric)
//   g[e][q] = (grad lambda)^T S (grad u)   at every quadrature point q of element e
// g is the derivative of lambda^T A(kappa) u with respect to kappa_q, so for a
// loss L(A(kappa) u) with lambda = dL/d(Au):  dL/dkappa = g, dL/du = v.
//
// Two thread layouts, as the forward kernel: the column layout
// (hex_fused.hpp: a Q x Q slab of threads per element with the z column in
// registers; u and lambda are interpolated and differentiated one after the
// other, the gradient of u kept in registers while lambda's is formed) and
// the line layout (hex_lines.hpp), chosen by the plan's layout.
#include "hex_plan.hpp"

namespace mi355 {
namespace {

template <int P, int Q>
__global__ void __launch_bounds__(256) hex_adjoint_kernel(int64_t elem_begin, int64_t elem_end, int scatter, const int32_t *__restrict__ offsets,
                                                          const int32_t *__restrict__ perm, const double *__restrict__ q_data,
                                                          const double *__restrict__ mats, const double *__restrict__ u,
                                                          const double *__restrict__ lam, double *__restrict__ v, double *__restrict__ evec,
                                                          const int32_t *__restrict__ tpos, double *__restrict__ g) {
  using S = HexShape<P, Q>;
  constexpr int T1 = S::T1, NT = S::NT, SLAB = S::SLAB, P3 = S::P3, Q3 = S::Q3, QC = 7;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *sB = reinterpret_cast<double *>(lds_raw);
  double *sD = sB + Q * P;
  const double *gB = mats, *gD = mats + Q * P;  // z-direction entries: wave-uniform scalar loads

  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  const int tx = t % T1, ty = t / T1;
  double   *S0 = sD + Q * Q + (size_t)le * 3 * SLAB;
  double   *S1 = S0 + SLAB;
  double   *S2 = S1 + SLAB;
  for (int i = threadIdx.x; i < Q * P + Q * Q; i += blockDim.x) sB[i] = mats[i];

  const int64_t e    = elem_begin + xcd_chunk(blockIdx.x, gridDim.x) * epb + le;
  const bool    live = le < epb && e < elem_end;
  const bool    node = live && tx < P && ty < P;
  auto sidx = [](int z, int y, int x) { return (z * T1 + y) * T1 + x; };

  int off[P];
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) off[z] = offsets[e * P3 + (z * P + ty) * P + tx];
  }
  const double *qd = q_data + (live ? e : 0) * (int64_t)QC * Q3 + (ty * Q + tx);
  __syncthreads();

  // values of `src` at the quadrature points of column (tx, ty), left in S0 as well
  auto interp = [&](const double *__restrict__ src, double (&xq)[Q]) {
    if (node) {
#pragma unroll
      for (int z = 0; z < P; z++) S0[sidx(z, ty, tx)] = src[off[z] >> 1];
    }
    __syncthreads();
    if (ty < P) {
      double acc[P];
#pragma unroll
      for (int z = 0; z < P; z++) acc[z] = 0;
#pragma unroll
      for (int b = 0; b < P; b++) {
        const double bb = sB[tx * P + b];
#pragma unroll
        for (int z = 0; z < P; z++) acc[z] += bb * S0[sidx(z, ty, b)];
      }
#pragma unroll
      for (int z = 0; z < P; z++) S1[sidx(z, ty, tx)] = acc[z];
    }
    __syncthreads();
    double acc[P];
#pragma unroll
    for (int z = 0; z < P; z++) acc[z] = 0;
#pragma unroll
    for (int b = 0; b < P; b++) {
      const double bb = sB[ty * P + b];
#pragma unroll
      for (int z = 0; z < P; z++) acc[z] += bb * S1[sidx(z, b, tx)];
    }
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      double s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += gB[qz * P + b] * acc[b];
      xq[qz] = s;
    }
    __syncthreads();  // S0 is rewritten next
#pragma unroll
    for (int qz = 0; qz < Q; qz++) S0[sidx(qz, ty, tx)] = xq[qz];
    __syncthreads();
  };
  // reference-space gradient at (tx, ty, qz) from the values in S0 and xq
  auto grad = [&](const double (&xq)[Q], int qz, double &dx, double &dy, double &dz) {
    dx = dy = dz = 0;
#pragma unroll
    for (int c = 0; c < Q; c++) {
      dx += sD[tx * Q + c] * S0[sidx(qz, ty, c)];
      dy += sD[ty * Q + c] * S0[sidx(qz, c, tx)];
      dz += gD[qz * Q + c] * xq[c];
    }
  };

  // ---- grad u, kept in registers
  double xq[Q], gux[Q], guy[Q], guz[Q];
  interp(u, xq);
#pragma unroll
  for (int qz = 0; qz < Q; qz++) grad(xq, qz, gux[qz], guy[qz], guz[qz]);
  __syncthreads();
  // ---- grad lambda; g = grad lambda . S grad u; w = kappa S grad lambda
  interp(lam, xq);
  double vz[Q];
#pragma unroll
  for (int qz = 0; qz < Q; qz++) {
    double lx, ly, lz;
    grad(xq, qz, lx, ly, lz);
    double wx = 0, wy = 0, wz = 0;
    if (live) {
      const double *q = qd + qz * Q * Q;
      poisson_point<double>(q, Q3, lx, ly, lz, wx, wy, wz);  // S grad lambda
      const double k = __builtin_nontemporal_load(q + 6 * Q3);
      g[(int64_t)perm[e] * Q3 + qz * Q * Q + ty * Q + tx] = wx * gux[qz] + wy * guy[qz] + wz * guz[qz];
      wx *= k;
      wy *= k;
      wz *= k;
    }
    S1[sidx(qz, ty, tx)] = wx;
    S2[sidx(qz, ty, tx)] = wy;
    vz[qz]               = wz;
  }
  __syncthreads();
  // ---- G^T then B^T (z in registers, y and x through LDS), as the forward kernel
  double acc[Q];
#pragma unroll
  for (int qz = 0; qz < Q; qz++) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      s += sD[a * Q + tx] * S1[sidx(qz, ty, a)];
      s += sD[a * Q + ty] * S2[sidx(qz, a, tx)];
      s += gD[a * Q + qz] * vz[a];
    }
    acc[qz] = s;
  }
  double t2[P];
#pragma unroll
  for (int b = 0; b < P; b++) {
    double s = 0;
#pragma unroll
    for (int qz = 0; qz < Q; qz++) s += gB[qz * P + b] * acc[qz];
    t2[b] = s;
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < P; b++) S0[sidx(b, ty, tx)] = t2[b];
  __syncthreads();
  if (ty < P) {
    double a2[P];
#pragma unroll
    for (int z = 0; z < P; z++) a2[z] = 0;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      const double bb = sB[a * P + ty];
#pragma unroll
      for (int z = 0; z < P; z++) a2[z] += bb * S0[sidx(z, a, tx)];
    }
#pragma unroll
    for (int z = 0; z < P; z++) S1[sidx(z, ty, tx)] = a2[z];
  }
  __syncthreads();
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) {
      double s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += sB[a * P + tx] * S1[sidx(z, ty, a)];
      if (scatter == kScatterEvec || (scatter == kScatterSplit && (off[z] & 1))) {  // L-entry-sorted E-vector
        evec[tpos[e * (int64_t)P3 + (z * P + ty) * P + tx]] = s;
        continue;
      }
      double *dst = &v[off[z] >> 1];
      if (scatter == kScatterSplit) {  // at most two contributions: onto the zeroed v
        unsafeAtomicAdd(dst, s);
        continue;
      }
      if (scatter == kScatterAtomic) unsafeAtomicAdd(dst, s);
      else if (scatter == kScatterOverwrite && (off[z] & 1)) *dst = s;
      else *dst += s;
    }
  }
}

// The same adjoint in the line thread layout (hex_lines.hpp): x and y
// contractions by whole 1D lines per thread, z contractions and the pointwise
// work on the (x, y) columns.  Five slabs per element: grad u's x and y parts
// stay in two of them while lambda is interpolated and differentiated through
// the other three.
template <int P, int Q>
__global__ void __launch_bounds__(256) hex_adjoint_lines_kernel(int64_t elem_begin, int64_t elem_end, int scatter, const int32_t *__restrict__ offsets,
                                                                const int32_t *__restrict__ perm, const double *__restrict__ q_data,
                                                                const double *__restrict__ mats, const double *__restrict__ u,
                                                                const double *__restrict__ lam, double *__restrict__ v, double *__restrict__ evec,
                                                                const int32_t *__restrict__ tpos, double *__restrict__ g) {
  constexpr int  NT = Q * Q, SLAB = Q * Q * Q, P3 = P * P * P, Q3 = Q * Q * Q, QC = 7, NS = 5;
  constexpr bool kRowWise = Q >= 7;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  double   *S0 = reinterpret_cast<double *>(lds_raw) + (size_t)le * NS * SLAB;
  double   *S1 = S0 + SLAB, *S2 = S1 + SLAB, *S3 = S2 + SLAB, *S4 = S3 + SLAB;
  auto      at = [](int z, int y, int x) { return (z * Q + y) * Q + x; };

  const int64_t e    = elem_begin + xcd_chunk(blockIdx.x, gridDim.x) * epb + le;
  const bool    live = le < epb && e < elem_end;
  const int     cx = t % Q, cy = t / Q, ny = t % P, nz = t / P;
  const bool    xline = t < P * P;
  const bool    node  = live && cx < P && cy < P;
  int           off[P];
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) off[z] = offsets[e * P3 + (z * P + cy) * P + cx];
  }
  const double *qd = q_data + (live ? e : 0) * (int64_t)QC * Q3 + (cy * Q + cx);
  double        sq[Q][QC];
#pragma unroll
  for (int qz = 0; qz < Q; qz++)
#pragma unroll
    for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);

  // reference-space gradient of src at the quadrature points: x part into SX,
  // y part into SY, z part in dz (this thread's column); Sa and Sb are
  // scratch (SY may be Sa: the last phase reads only Sb)
  auto gradient = [&](const double *__restrict__ src, double *SX, double *SY, double *Sa, double *Sb, double(&dz)[Q]) {
    if (node) {
#pragma unroll
      for (int z = 0; z < P; z++) Sa[at(z, cy, cx)] = src[off[z] >> 1];
    }
    __syncthreads();
    if (xline) {  // B_x on the x-lines of nodes
      const auto gB = phase_ptr(mats);
      double     in[P];
#pragma unroll
      for (int x = 0; x < P; x++) in[x] = Sa[at(nz, ny, x)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? phase_ptr(mats + qx * P) : gB + qx * P;
        double     s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * in[b];
        Sb[at(nz, ny, qx)] = s;
      }
    }
    __syncthreads();
    if (t < Q * P) {  // B_y on the y-lines (qx, z), z < P
      const auto gB = phase_ptr(mats);
      const int  qx = t % Q, z = t / Q;
      double     in[P];
#pragma unroll
      for (int b = 0; b < P; b++) in[b] = Sb[at(z, b, qx)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? phase_ptr(mats + qy * P) : gB + qy * P;
        double     s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * in[b];
        Sa[at(z, qy, qx)] = s;
      }
    }
    __syncthreads();
    {  // B_z and D_z on the column; the values go to Sb for the x/y derivatives
      const auto gB = phase_ptr(mats);
      double     col[P], xq[Q];
#pragma unroll
      for (int b = 0; b < P; b++) col[b] = Sa[at(b, cy, cx)];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + qz * P) : gB + qz * P;
        double     s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * col[b];
        xq[qz] = s;
      }
      const auto gD = phase_ptr(mats + Q * P);
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qz * Q) : gD + qz * Q;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * xq[c];
        dz[qz]              = s;
        Sb[at(qz, cy, cx)] = xq[qz];
      }
    }
    __syncthreads();
    {  // D_x by x-lines (qy = a, qz) into SX, D_y by y-lines (qx = a, qz) into SY
      const auto gD = phase_ptr(mats + Q * P);
      const int  a = t % Q, qz = t / Q;
      double     in[Q];
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = Sb[at(qz, a, c)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qx * Q) : gD + qx * Q;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        SX[at(qz, a, qx)] = s;
      }
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = Sb[at(qz, c, a)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qy * Q) : gD + qy * Q;
        double     s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        SY[at(qz, qy, a)] = s;
      }
    }
    __syncthreads();
  };

  double dzu[Q], dzl[Q];
  gradient(u, S2, S3, S0, S1, dzu);    // grad u: S2 (x), S3 (y), dzu (z)
  gradient(lam, S4, S0, S0, S1, dzl);  // grad lambda: S4 (x), S0 (y), dzl (z)
  // ---- pointwise: w = S grad lambda, g = w . grad u, then kappa w (in place on the column)
  double vz[Q];
#pragma unroll
  for (int qz = 0; qz < Q; qz++) {
    double wx, wy, wz;
    poisson_regs<double>(sq[qz], S4[at(qz, cy, cx)], S0[at(qz, cy, cx)], dzl[qz], wx, wy, wz);
    if (live) g[(int64_t)perm[e] * Q3 + qz * Q * Q + cy * Q + cx] = wx * S2[at(qz, cy, cx)] + wy * S3[at(qz, cy, cx)] + wz * dzu[qz];
    const double k = live ? sq[qz][6] : 0.0;
    S4[at(qz, cy, cx)] = wx * k;
    S0[at(qz, cy, cx)] = wy * k;
    vz[qz]             = wz * k;
  }
  __syncthreads();
  {  // D_x^T by x-lines into S2, D_y^T by y-lines into S3
    const auto gD = phase_ptr(mats + Q * P);
    const int  a = t % Q, qz = t / Q;
    double     in[Q];
#pragma unroll
    for (int c = 0; c < Q; c++) in[c] = S4[at(qz, a, c)];
#pragma unroll
    for (int qx = 0; qx < Q; qx++) {
      const auto m = kRowWise ? phase_ptr(mats + Q * P + qx) : gD + qx;
      double     s = 0;
#pragma unroll
      for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
      S2[at(qz, a, qx)] = s;
    }
#pragma unroll
    for (int c = 0; c < Q; c++) in[c] = S0[at(qz, c, a)];
#pragma unroll
    for (int qy = 0; qy < Q; qy++) {
      const auto m = kRowWise ? phase_ptr(mats + Q * P + qy) : gD + qy;
      double     s = 0;
#pragma unroll
      for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
      S3[at(qz, qy, a)] = s;
    }
  }
  __syncthreads();
  {  // sum of the three transposes on the column, then B_z^T into S1
    const auto gDz = phase_ptr(mats + Q * P);
    double     w[Q];
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      const auto m = kRowWise ? phase_ptr(mats + Q * P + qz) : gDz + qz;
      double     s = S2[at(qz, cy, cx)] + S3[at(qz, cy, cx)];
#pragma unroll
      for (int c = 0; c < Q; c++) s += m[c * Q] * vz[c];
      w[qz] = s;
    }
    const auto gBz = phase_ptr(mats);
#pragma unroll
    for (int b = 0; b < P; b++) {
      const auto m = kRowWise ? phase_ptr(mats + b) : gBz + b;
      double     s = 0;
#pragma unroll
      for (int qz = 0; qz < Q; qz++) s += m[qz * P] * w[qz];
      S1[at(b, cy, cx)] = s;
    }
  }
  __syncthreads();
  if (t < Q * P) {  // B_y^T on the y-lines (qx, z), z < P, into S0
    const auto gB = phase_ptr(mats);
    const int  qx = t % Q, z = t / Q;
    double     in[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) in[a] = S1[at(z, a, qx)];
#pragma unroll
    for (int y = 0; y < P; y++) {
      const auto m = kRowWise ? phase_ptr(mats + y) : gB + y;
      double     s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
      S0[at(z, y, qx)] = s;
    }
  }
  __syncthreads();
  if (xline) {  // B_x^T on the x-lines of nodes, into S1
    const auto gB = phase_ptr(mats);
    double     in[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) in[a] = S0[at(nz, ny, a)];
#pragma unroll
    for (int x = 0; x < P; x++) {
      const auto m = kRowWise ? phase_ptr(mats + x) : gB + x;
      double     s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
      S1[at(nz, ny, x)] = s;
    }
  }
  __syncthreads();
  if (node) {  // E^T from the node columns
    double s[P];
#pragma unroll
    for (int z = 0; z < P; z++) s[z] = S1[at(z, cy, cx)];
    if (scatter == kScatterEvec) {
#pragma unroll
      for (int z = 0; z < P; z++) evec[tpos[e * (int64_t)P3 + (z * P + cy) * P + cx]] = s[z];
    } else if (scatter == kScatterSplit) {  // shared entries to their E-vector slot, the rest onto the zeroed v
#pragma unroll
      for (int z = 0; z < P; z++) {
        if (off[z] & 1) evec[tpos[e * (int64_t)P3 + (z * P + cy) * P + cx]] = s[z];
        else unsafeAtomicAdd(&v[off[z] >> 1], s[z]);
      }
    } else if (scatter == kScatterAtomic) {
#pragma unroll
      for (int z = 0; z < P; z++) unsafeAtomicAdd(&v[off[z] >> 1], s[z]);
    } else {  // colour-exclusive; first element to reach a node writes it
#pragma unroll
      for (int z = 0; z < P; z++) {
        double *dst = &v[off[z] >> 1];
        *dst        = (off[z] & 1) ? s[z] : *dst + s[z];
      }
    }
  }
}

__global__ void __launch_bounds__(256) zero_fill(int64_t n, double *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = 0.0;
}

template <int P, int Q>
int launch_adjoint(const CeedMi355Operator_private *op, const double *u, const double *lam, double *v, double *g, hipStream_t s) {
  const int    nt    = Q * Q;
  const int    epb   = op->epb;
  const bool   lines = op->layout == kLayoutLines;  // (plans in the point layout take the column kernel here)
  const size_t lds   = sizeof(double) * (size_t)(lines ? epb * 5 * Q * Q * Q : Q * P + Q * Q + epb * 3 * Q * Q * Q);
  const bool   tmap  = op->scatter == 0;
  MI355_CHECK(op->scatter != CEED_MI355_PLAN_SEGMENT, CEED_MI355_ERROR_UNSUPPORTED,
              "the adjoint does not take segment plans (elements of a colour share segment faces)");
  MI355_CHECK(lds <= (size_t)kMaxLdsBytes, CEED_MI355_ERROR_BACKEND, "adjoint needs %zu bytes of LDS at %d elements per workgroup", lds, epb);
  for (size_t c = 0; c + 1 < op->color_start.size(); c++) {
    const int64_t b = op->color_start[c], e = op->color_start[c + 1];
    if (e <= b) continue;
    const int64_t grid    = (e - b + epb - 1) / epb;
    const int     scatter = tmap ? kScatterEvec
                                 : op->scatter == CEED_MI355_PLAN_SPLIT ? kScatterSplit
                                                                        : (op->scatter == CEED_MI355_PLAN_ATOMIC ? kScatterAtomic : kScatterOverwrite);
    if (lines)
      hex_adjoint_lines_kernel<P, Q><<<(unsigned)grid, epb * nt, lds, s>>>(b, e, scatter, op->d_offsets, op->d_perm, (const double *)op->d_qdata,
                                                                          (const double *)op->d_mats, u, lam, v, (double *)op->d_evec, op->d_tpos, g);
    else
      hex_adjoint_kernel<P, Q><<<(unsigned)grid, epb * nt, lds, s>>>(b, e, scatter, op->d_offsets, op->d_perm, (const double *)op->d_qdata,
                                                                    (const double *)op->d_mats, u, lam, v, (double *)op->d_evec, op->d_tpos, g);
  }
  MI355_LAUNCH_CHECK();
  if (tmap) return transpose_gather(op, v, 1, s);
  if (op->scatter == CEED_MI355_PLAN_SPLIT) return split_finish(op, v, 1, s);
  return CEED_MI355_SUCCESS;
}

using AdjointFn = int (*)(const CeedMi355Operator_private *, const double *, const double *, double *, double *, hipStream_t);

AdjointFn pick_adjoint(int P, int Q) {
#define MI355_ADJ_CASE(PP, QQ) \
  if (P == PP && Q == QQ) return launch_adjoint<PP, QQ>;
  MI355_ADJ_CASE(2, 2) MI355_ADJ_CASE(2, 3) MI355_ADJ_CASE(3, 3) MI355_ADJ_CASE(3, 4) MI355_ADJ_CASE(4, 4) MI355_ADJ_CASE(4, 5)
  MI355_ADJ_CASE(5, 5) MI355_ADJ_CASE(5, 6) MI355_ADJ_CASE(6, 6) MI355_ADJ_CASE(6, 7) MI355_ADJ_CASE(7, 7) MI355_ADJ_CASE(7, 8)
  MI355_ADJ_CASE(8, 8) MI355_ADJ_CASE(8, 9) MI355_ADJ_CASE(9, 9) MI355_ADJ_CASE(9, 10)
#undef MI355_ADJ_CASE
  return nullptr;
}

}  // namespace
}  // namespace mi355

extern "C" int CeedMi355OperatorApplyAdjoint(CeedMi355Operator op, const double *u, const double *lambda, double *v, double *dkappa,
                                             void *stream) {
  MI355_CHECK(op, CEED_MI355_ERROR_MAJOR, "null operator");
  MI355_CHECK(op->qf == mi355::kQfDiffusionKappa && op->scalar_type == CEED_MI355_FP64 && op->num_comp == 1, CEED_MI355_ERROR_BACKEND,
              "the adjoint needs an FP64 scalar kappa-diffusion plan");
  hipStream_t s = (hipStream_t)stream;
  if (op->num_elem == 0) return CEED_MI355_SUCCESS;
  auto fn = mi355::pick_adjoint(op->P, op->Q);
  MI355_CHECK(fn, CEED_MI355_ERROR_UNSUPPORTED, "no adjoint kernel for P=%d Q=%d", op->P, op->Q);
  if (!op->covers_all) {  // a fill kernel, not a memset node: safe under graph capture
    mi355::zero_fill<<<mi355_blocks(op->l_size, 256, 4096), 256, 0, s>>>(op->l_size, v);
    MI355_LAUNCH_CHECK();
  }
  return fn(op, u, lambda, v, dkappa, s);
}
