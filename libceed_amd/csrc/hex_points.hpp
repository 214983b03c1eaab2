// Fused per-element hex operator, "point" thread layout (a 3D slab).
//
// Same operator as hex_apply_kernel and hex_apply_lines_kernel (gather E,
// sum-factorised interp and collocated gradient B, pointwise D, transposes,
// E^T), with one thread per quadrature point of the element: every
// contraction, z included, goes through LDS, each thread forming the one
// output value of its point from the Q (or P) inputs along the contracted
// direction.  Q^3 threads per element (Q <= 6 keeps an element within a
// 256-thread workgroup); B and D staged in LDS once per workgroup, then four
// slabs of Q^3 values per element (three for mass; the interpolated values
// stay readable while the three flux components are written).  Every sum
// runs over the same terms in the same order as in the other two layouts
// (the pointwise operator is shared), so the three give bitwise identical
// results and the autotuner can pick any of them.
//
// Thread t = x + Q (y + Q z) of an element plays:
//   point (x, y, z)                      B_z, D, transposes
//   node (x, y, z), x, y, z < P          gather E, scatter E^T
//   (x, y, z), y, z < P                  B_x, B_x^T
//   (x, y, z), z < P                     B_y, B_y^T
#pragma once

#include "hex_fused.hpp"

namespace mi355 {

template <int QF>
struct PointSlabs {
  static constexpr int kSlabs = QF == kQfMass ? 3 : 4;
};

template <int P, int Q, int QF>
__host__ __device__ constexpr int hex_points_lds_entries(int epb) {
  return Q * P + Q * Q + epb * PointSlabs<QF>::kSlabs * Q * Q * Q;
}

template <int P, int Q, int NCOMP, int QF, typename T>
__global__ void __launch_bounds__(256) hex_apply_points_kernel(int64_t elem_begin, int64_t elem_end, int scatter,
                                                                const int32_t *__restrict__ offsets, int32_t comp_stride,
                                                                const T *__restrict__ q_data, const T *__restrict__ mats,
                                                                const T *__restrict__ u, T *__restrict__ v, T *__restrict__ evec,
                                                                const int32_t *__restrict__ tpos, const int32_t *__restrict__ elem_sys,
                                                                const int32_t *__restrict__ sys_active) {
  constexpr int NT = Q * Q * Q, SLAB = Q * Q * Q, P3 = P * P * P, QC = QfTraits<QF>::kComps;
  static_assert(NT <= 256, "the point layout needs Q^3 <= 256");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  // B (Q x P) then D (Q x Q), row-major, at mats, staged in LDS
  T *gB = reinterpret_cast<T *>(lds_raw);
  T *gD = gB + Q * P;
  for (int i = threadIdx.x; i < Q * P + Q * Q; i += blockDim.x) gB[i] = mats[i];

  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  const int x = t % Q, y = (t / Q) % Q, z = t / (Q * Q);
  T        *S0 = gD + Q * Q + (size_t)le * PointSlabs<QF>::kSlabs * SLAB;
  T        *S1 = S0 + SLAB, *S2 = S1 + SLAB;
  T        *S3 = S2 + SLAB;  // diffusion only
  auto      at = [](int zz, int yy, int xx) { return (zz * Q + yy) * Q + xx; };

  const int64_t e    = elem_begin + xcd_chunk(blockIdx.x, gridDim.x) * epb + le;
  const bool    live = le < epb && e < elem_end && (!elem_sys || sys_active[elem_sys[e]]);
  if (elem_sys && !__syncthreads_or(live)) return;  // every element of the workgroup is frozen (uniform exit)
  const bool node = live && x < P && y < P && z < P;

  int off = 0;  // 2 * node + first_touch
  if (node) off = offsets[e * P3 + (z * P + y) * P + x];
  // the point's geometric factors (dead lanes read element 0's, unused)
  T sq[QC];
  {
    const T *qd = q_data + (live ? e : 0) * (int64_t)QC * SLAB + t;
#pragma unroll
    for (int k = 0; k < QC; k++) sq[k] = __builtin_nontemporal_load(&qd[k * SLAB]);
  }

  for (int comp = 0; comp < NCOMP; comp++) {
    const int64_t coff = (int64_t)comp * comp_stride;
    // ---- E: S0[z][y][x] = u at the node (x, y, z < P); the previous component
    // last read S0 before the barrier ahead of its B_x^T
    // (a component's B_x^T reads S1 (mass) or S2, which the next component
    // writes only after the barrier behind E)
    if (node) S0[at(z, y, x)] = u[(off >> 1) + coff];
    __syncthreads();
    // ---- B_x: S1[z][y][x] = sum_b B[x][b] S0[z][y][b]   (y, z < P)
    if (live && y < P && z < P) {
      T s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += gB[x * P + b] * S0[at(z, y, b)];
      S1[at(z, y, x)] = s;
    }
    __syncthreads();
    // ---- B_y: S2[z][y][x] = sum_b B[y][b] S1[z][b][x]   (z < P)
    if (live && z < P) {
      T s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += gB[y * P + b] * S1[at(z, b, x)];
      S2[at(z, y, x)] = s;
    }
    __syncthreads();
    // ---- B_z: the interpolated value at the point
    T uq = 0;
#pragma unroll
    for (int b = 0; b < P; b++) uq += gB[z * P + b] * S2[at(b, y, x)];
    if constexpr (QF == kQfMass) {
      // ---- D, then B_z^T into S0 (z < P) after every thread has read S2
      S1[at(z, y, x)] = uq * (live ? sq[0] : T(0));
      __syncthreads();
      if (live && z < P) {
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += gB[qz * P + z] * S1[at(qz, y, x)];
        S0[at(z, y, x)] = s;
      }
    } else {
      // ---- collocated gradient at the point, the pointwise operator, and the
      // three flux components into S1, S2, S3
      S0[at(z, y, x)] = uq;
      __syncthreads();
      T dx = 0, dy = 0, dz = 0;
#pragma unroll
      for (int c = 0; c < Q; c++) {
        dx += gD[x * Q + c] * S0[at(z, y, c)];
        dy += gD[y * Q + c] * S0[at(z, c, x)];
        dz += gD[z * Q + c] * S0[at(c, y, x)];
      }
      T vx, vy, vzz;
      poisson_regs<T>(sq, dx, dy, dz, vx, vy, vzz);
      if constexpr (QF == kQfDiffusionKappa) {
        vx *= sq[6];
        vy *= sq[6];
        vzz *= sq[6];
      }
      if (!live) vx = vy = vzz = 0;
      S1[at(z, y, x)] = vx;  // (S1 and S2 were last read before the barrier above)
      S2[at(z, y, x)] = vy;
      S3[at(z, y, x)] = vzz;
      __syncthreads();
      // ---- the three transposes: x and y sums first, then the z terms
      T sx = 0, sy = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) {
        sx += gD[a * Q + x] * S1[at(z, y, a)];
        sy += gD[a * Q + y] * S2[at(z, a, x)];
      }
      T w = sx + sy;
#pragma unroll
      for (int a = 0; a < Q; a++) w += gD[a * Q + z] * S3[at(a, y, x)];
      S0[at(z, y, x)] = w;
      __syncthreads();
      // ---- B_z^T into S1 (z < P)
      if (live && z < P) {
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += gB[qz * P + z] * S0[at(qz, y, x)];
        S1[at(z, y, x)] = s;
      }
    }
    __syncthreads();
    T *Sz = QF == kQfMass ? S0 : S1;  // the B_z^T result
    T *Sy = QF == kQfMass ? S1 : S2;
    // ---- B_y^T: Sy[z][y][x] = sum_a B[a][y] Sz[z][a][x]   (y, z < P)
    if (live && y < P && z < P) {
      T s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += gB[a * P + y] * Sz[at(z, a, x)];
      Sy[at(z, y, x)] = s;
    }
    __syncthreads();
    // ---- B_x^T at the node, then E^T
    if (node) {
      T s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += gB[a * P + x] * Sy[at(z, y, a)];
      if (scatter == kScatterEvec) {  // tpos gives each contribution's slot in the sorted E-vector
        evec[tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + y) * P + x]] = s;
      } else if (scatter == kScatterAtomic) {
        unsafeAtomicAdd(&v[(off >> 1) + coff], s);
      } else if (scatter == kScatterSplit) {  // shared entries to their E-vector slot, the rest onto the zeroed target
        if (off & 1) evec[tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + y) * P + x]] = s;
        else unsafeAtomicAdd(&v[(off >> 1) + coff], s);
      } else {
        T *dst = &v[(off >> 1) + coff];
        *dst   = (scatter == kScatterOverwrite && (off & 1)) ? s : *dst + s;
      }
    }
  }
}

}  // namespace mi355
