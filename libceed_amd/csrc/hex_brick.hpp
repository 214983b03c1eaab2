// Fused per-element hex operator, "brick" E^T.
//
// The plan groups the elements into bricks of up to bx * by * bz face-sharing
// elements (2 x 2 x 2 on a structured mesh at p = 4).  One workgroup owns one
// brick and walks its elements one after another, each element through the
// same line-layout pipeline as hex_apply_lines_kernel (gather E, interp and
// collocated gradient B, pointwise D, transposes).  Instead of a global E^T
// per element, every element adds its P^3 results into an LDS accumulator
// over the brick's nodes, in the brick's slot order, so the nodes shared
// inside the brick are summed on chip.  At the end of the brick:
//   * nodes whose every element lies in the brick ("complete") are written to
//     v once (Apply) or added to it (ApplyAdd): coalesced, no atomics;
//   * the other nodes (on the brick's surface) store their partial sum into a
//     per-brick slot, and brick_gather_kernel adds the partials of each such
//     node in ascending brick order.
// Every sum has a fixed order (slot order inside a brick, brick order across
// bricks), so the result is bitwise reproducible; the order differs from the
// reference's ascending (elem, comp, node) order only by rounding.
//
// The brick's node list (2 * L-node + incomplete bit, or -1 where no element
// of the brick lies) is staged in LDS once; an element's gather reads its
// node indices from it by the brick-local position, so the only global load
// on the gather's path is u itself.  The next element's u values and q_data
// are loaded right after the current element's pointwise operator has used
// its q_data: their latency overlaps the back half of the current element,
// and the registers they take are the ones the current q_data has just freed.
//
// Thread t of the workgroup's Q*Q threads plays the roles of the line layout:
//   column (qx, qy) = (t % Q, t / Q)                     z-contractions, pointwise D
//   node column (x, y) = (t % Q, t / Q), x, y < P      gather E, LDS accumulation
//   x-line of nodes (y, z) = (t % P, t / P), t < P*P     B_x, B_x^T
//   y-line (qx, z) = (t % Q, t / Q), t < Q*P              B_y, B_y^T
//   x-line (qy, qz) and y-line (qx, qz) = (t % Q, t / Q)  the collocated derivatives
#pragma once

#include <type_traits>

#include "hex_lines.hpp"

namespace mi355 {

// LDS bytes of one brick workgroup: the element's line-layout slabs, the
// accumulator (num_comp values per brick node) and the node list.
template <int P, int Q, int NCOMP, int QF, typename T>
__host__ __device__ constexpr size_t hex_brick_lds_bytes(int nb_nodes) {
  return sizeof(T) * ((size_t)LineSlabs<QF>::kSlabs * Q * Q * Q + (size_t)NCOMP * nb_nodes) + sizeof(int32_t) * (size_t)nb_nodes;
}

// PREFETCH: issue the next element's loads during the current one (above), or
// each element's own loads at its start, as the line kernel does (fewer live
// registers: 127 against 137 VGPRs at p = 4, FP64 diffusion).
template <int P, int Q, int NCOMP, int QF, typename T, bool PREFETCH>
__global__ void __launch_bounds__(256) hex_apply_brick_kernel(int bx, int by, int nb_nodes, const int32_t *__restrict__ brick_start,
                                                               const int32_t *__restrict__ brick_mask, const int32_t *__restrict__ bnode,
                                                               int32_t comp_stride, const T *__restrict__ q_data,
                                                               const T *__restrict__ mats, const T *__restrict__ u, T *__restrict__ v,
                                                               T *__restrict__ partials, int overwrite) {
  constexpr int NT = Q * Q, SLAB = Q * Q * Q, Q3 = Q * Q * Q, QC = QfTraits<QF>::kComps;
  constexpr int NS = LineSlabs<QF>::kSlabs, p = P - 1;
  constexpr bool kRowWise = Q >= 7;  // as in the line kernel: matrix entries per output row from Q = 7
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T       *S0  = reinterpret_cast<T *>(lds_raw);
  T       *S1  = S0 + SLAB;
  T       *S2  = S1 + SLAB;  // diffusion only
  T       *ACC = S0 + NS * SLAB;                                  // [NCOMP][nb_nodes]
  int32_t *sN  = reinterpret_cast<int32_t *>(ACC + NCOMP * nb_nodes);  // [nb_nodes]
  auto     at  = [](int z, int y, int x) { return (z * Q + y) * Q + x; };

  const int     t = threadIdx.x;
  const int64_t b = xcd_chunk(blockIdx.x, gridDim.x);  // grid = number of bricks
  for (int i = t; i < nb_nodes; i += NT) {
    sN[i] = bnode[b * nb_nodes + i];
#pragma unroll
    for (int c = 0; c < NCOMP; c++) ACC[c * nb_nodes + i] = T(0);
  }
  const int64_t e0   = brick_start[b];
  const int     mask = brick_mask[b];
  const int     nbx = bx * p + 1, nbxy = nbx * (by * p + 1);
  const int     cx = t % Q, cy = t / Q;  // column; node column when both < P
  const int     ny = t % P, nz = t / P;  // x-line of nodes
  const bool    xline = t < P * P;
  const bool    node  = cx < P && cy < P;
  // brick-local index of node (cx, cy, 0) of the element in slot s; +z*nbxy for node z
  auto local0 = [&](int s) {
    const int sx = s % bx, sy = (s / bx) % by, sz = s / (bx * by);
    return (sz * p) * nbxy + (sy * p + cy) * nbx + sx * p + cx;
  };
  T    ug[NCOMP][P];
  T    sq[Q][QC];
  auto gather = [&](int s) {
    if (node) {
      const int l0 = local0(s);
#pragma unroll
      for (int z = 0; z < P; z++) {
        const int64_t n = sN[l0 + z * nbxy] >> 1;
#pragma unroll
        for (int c = 0; c < NCOMP; c++) ug[c][z] = u[n + (int64_t)c * comp_stride];
      }
    }
  };
  auto load_q = [&](int64_t e) {
    const T *qd = q_data + e * (int64_t)QC * Q3 + (cy * Q + cx);
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
#pragma unroll
      for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
  };
  // one element of the brick; with kPrefetch the next element's u and q_data
  // are issued after this one's pointwise operator (a separate copy without
  // the loads for the last element, so that the loads are unconditional in
  // the copy that has them and take over the registers of the q_data just used)
  auto element = [&](auto prefetch_tag, int s, int s_next, int64_t e) {
    constexpr bool kPrefetch = PREFETCH && decltype(prefetch_tag)::value;
    if constexpr (!PREFETCH) {
      gather(s);
      load_q(e);
    }
    for (int comp = 0; comp < NCOMP; comp++) {
      // ---- E: the element's nodes (gathered into registers beforehand) into S0
      if (node) {
#pragma unroll
        for (int z = 0; z < P; z++) S0[at(z, cy, cx)] = ug[comp][z];
      }
      __syncthreads();
      // ---- B_x: x-line (ny, nz) of nodes, Q interpolated values into S1
      if (xline) {
        const auto gB = phase_ptr(mats);
        T          in[P];
#pragma unroll
        for (int x = 0; x < P; x++) in[x] = S0[at(nz, ny, x)];
#pragma unroll
        for (int qx = 0; qx < Q; qx++) {
          const auto m = kRowWise ? phase_ptr(mats + qx * P) : gB + qx * P;
          T          acc = 0;
#pragma unroll
          for (int i = 0; i < P; i++) acc += m[i] * in[i];
          S1[at(nz, ny, qx)] = acc;
        }
      }
      __syncthreads();
      // ---- B_y: y-line (qx, z), z < P
      if (t < Q * P) {
        const auto gB = phase_ptr(mats);
        const int  qx = t % Q, z = t / Q;
        T          in[P];
#pragma unroll
        for (int i = 0; i < P; i++) in[i] = S1[at(z, i, qx)];
#pragma unroll
        for (int qy = 0; qy < Q; qy++) {
          const auto m = kRowWise ? phase_ptr(mats + qy * P) : gB + qy * P;
          T          acc = 0;
#pragma unroll
          for (int i = 0; i < P; i++) acc += m[i] * in[i];
          S0[at(z, qy, qx)] = acc;
        }
      }
      __syncthreads();
      // ---- B_z on the column (cx, cy): uq[qz]
      T uq[Q];
      {
        const auto gB = phase_ptr(mats);
        T          col[P];
#pragma unroll
        for (int i = 0; i < P; i++) col[i] = S0[at(i, cy, cx)];
#pragma unroll
        for (int qz = 0; qz < Q; qz++) {
          const auto m = kRowWise ? phase_ptr(mats + qz * P) : gB + qz * P;
          T          acc = 0;
#pragma unroll
          for (int i = 0; i < P; i++) acc += m[i] * col[i];
          uq[qz] = acc;
        }
      }
      const bool prefetch = kPrefetch && comp == NCOMP - 1;
      if constexpr (QF == kQfMass) {
        // ---- D, then B_z^T back into the column's own S0 entries
#pragma unroll
        for (int qz = 0; qz < Q; qz++) uq[qz] *= sq[qz][0];
        if (prefetch) {
          gather(s_next);
          load_q(e + 1);
        }
        const auto gB = phase_ptr(mats);
#pragma unroll
        for (int i = 0; i < P; i++) {
          const auto m = kRowWise ? phase_ptr(mats + i) : gB + i;
          T          acc = 0;
#pragma unroll
          for (int qz = 0; qz < Q; qz++) acc += m[qz * P] * uq[qz];
          S0[at(i, cy, cx)] = acc;
        }
      } else {
        // ---- collocated derivatives: z on the column, x and y by lines through S1
        const auto gD = phase_ptr(mats + Q * P);
        T          dz[Q];
#pragma unroll
        for (int qz = 0; qz < Q; qz++) {
          const auto m = kRowWise ? phase_ptr(mats + Q * P + qz * Q) : gD + qz * Q;
          T          acc = 0;
#pragma unroll
          for (int c = 0; c < Q; c++) acc += m[c] * uq[c];
          dz[qz]             = acc;
          S1[at(qz, cy, cx)] = uq[qz];
        }
        __syncthreads();
        {
          const auto gD = phase_ptr(mats + Q * P);
          const int  a = t % Q, qz = t / Q;  // x-line (qy = a, qz) and y-line (qx = a, qz)
          T          in[Q];
#pragma unroll
          for (int c = 0; c < Q; c++) in[c] = S1[at(qz, a, c)];
#pragma unroll
          for (int qx = 0; qx < Q; qx++) {
            const auto m = kRowWise ? phase_ptr(mats + Q * P + qx * Q) : gD + qx * Q;
            T          acc = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) acc += m[c] * in[c];
            S2[at(qz, a, qx)] = acc;
          }
#pragma unroll
          for (int c = 0; c < Q; c++) in[c] = S1[at(qz, c, a)];
#pragma unroll
          for (int qy = 0; qy < Q; qy++) {
            const auto m = kRowWise ? phase_ptr(mats + Q * P + qy * Q) : gD + qy * Q;
            T          acc = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) acc += m[c] * in[c];
            S0[at(qz, qy, a)] = acc;
          }
        }
        __syncthreads();
        // ---- pointwise (Poisson3DApply, times kappa for the kappa-weighted operator) on the column
        T vz[Q];
#pragma unroll
        for (int qz = 0; qz < Q; qz++) {
          T vx, vy, vzz;
          poisson_regs<T>(sq[qz], S2[at(qz, cy, cx)], S0[at(qz, cy, cx)], dz[qz], vx, vy, vzz);
          if constexpr (QF == kQfDiffusionKappa) {
            vx *= sq[qz][6];
            vy *= sq[qz][6];
            vzz *= sq[qz][6];
          }
          S2[at(qz, cy, cx)] = vx;
          S0[at(qz, cy, cx)] = vy;
          vz[qz]             = vzz;
        }
        // the next element's u and q_data, in flight during the rest of this one
        // (kept below the pointwise operator: issued earlier, they would hold
        // their registers while this element's q_data are still live)
        if (prefetch) {
          gather(s_next);
          load_q(e + 1);
        }
        __syncthreads();
        // ---- D_x^T by x-lines into S1; D_y^T by y-lines, held in registers
        // until every x-line has read S2, then written over it
        {
          const int a = t % Q, qz = t / Q;
          T         ty[Q];
          {
            const auto gD = phase_ptr(mats + Q * P);
            T          in[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) in[c] = S2[at(qz, a, c)];
#pragma unroll
            for (int qx = 0; qx < Q; qx++) {
              const auto m = kRowWise ? phase_ptr(mats + Q * P + qx) : gD + qx;
              T          acc = 0;
#pragma unroll
              for (int c = 0; c < Q; c++) acc += m[c * Q] * in[c];
              S1[at(qz, a, qx)] = acc;
            }
          }
          {
            const auto gD = phase_ptr(mats + Q * P);
            T          in[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) in[c] = S0[at(qz, c, a)];
#pragma unroll
            for (int qy = 0; qy < Q; qy++) {
              const auto m = kRowWise ? phase_ptr(mats + Q * P + qy) : gD + qy;
              T          acc = 0;
#pragma unroll
              for (int c = 0; c < Q; c++) acc += m[c * Q] * in[c];
              ty[qy] = acc;
            }
          }
          __syncthreads();
#pragma unroll
          for (int qy = 0; qy < Q; qy++) S2[at(qz, qy, a)] = ty[qy];
        }
        __syncthreads();
        // ---- sum of the three transposes on the column, then B_z^T into S0
        const auto gDz = phase_ptr(mats + Q * P);
        T          w[Q];
#pragma unroll
        for (int qz = 0; qz < Q; qz++) {
          const auto m   = kRowWise ? phase_ptr(mats + Q * P + qz) : gDz + qz;
          T          acc = S1[at(qz, cy, cx)] + S2[at(qz, cy, cx)];
#pragma unroll
          for (int c = 0; c < Q; c++) acc += m[c * Q] * vz[c];
          w[qz] = acc;
        }
        const auto gBz = phase_ptr(mats);
#pragma unroll
        for (int i = 0; i < P; i++) {
          const auto m   = kRowWise ? phase_ptr(mats + i) : gBz + i;
          T          acc = 0;
#pragma unroll
          for (int qz = 0; qz < Q; qz++) acc += m[qz * P] * w[qz];
          S0[at(i, cy, cx)] = acc;
        }
      }
      __syncthreads();
      // ---- B_y^T: y-line (qx, z), z < P, into S1
      if (t < Q * P) {
        const auto gB = phase_ptr(mats);
        const int  qx = t % Q, z = t / Q;
        T          in[Q];
#pragma unroll
        for (int a = 0; a < Q; a++) in[a] = S0[at(z, a, qx)];
#pragma unroll
        for (int y = 0; y < P; y++) {
          const auto m   = kRowWise ? phase_ptr(mats + y) : gB + y;
          T          acc = 0;
#pragma unroll
          for (int a = 0; a < Q; a++) acc += m[a * P] * in[a];
          S1[at(z, y, qx)] = acc;
        }
      }
      __syncthreads();
      // ---- B_x^T: x-line (ny, nz) of nodes, into S0
      if (xline) {
        const auto gB = phase_ptr(mats);
        T          in[Q];
#pragma unroll
        for (int a = 0; a < Q; a++) in[a] = S1[at(nz, ny, a)];
#pragma unroll
        for (int x = 0; x < P; x++) {
          const auto m   = kRowWise ? phase_ptr(mats + x) : gB + x;
          T          acc = 0;
#pragma unroll
          for (int a = 0; a < Q; a++) acc += m[a * P] * in[a];
          S0[at(nz, ny, x)] = acc;
        }
      }
      __syncthreads();
      // ---- E^T into the brick accumulator, from the node columns (distinct
      // nodes of one element never share a brick position, so no two threads
      // add to the same entry; elements follow one another behind barriers)
      if (node) {
        T        *acc = ACC + comp * nb_nodes + local0(s);
#pragma unroll
        for (int z = 0; z < P; z++) acc[z * nbxy] += S0[at(z, cy, cx)];
      }
    }
  };
  __syncthreads();  // node list and zeroed accumulator
  int     s = __builtin_ctz(mask);
  int64_t e = e0;
  if constexpr (PREFETCH) {
    gather(s);
    load_q(e);
  }
  for (int rest = mask & ~((2 << s) - 1); rest; rest &= rest - 1) {
    const int s_next = __builtin_ctz(rest);
    element(std::true_type{}, s, s_next, e);
    s = s_next;
    e++;
  }
  element(std::false_type{}, s, -1, e);
  __syncthreads();
  // ---- the brick's nodes: complete ones to v, the others to their partial slot
  for (int i = t; i < nb_nodes; i += NT) {
    const int32_t g2 = sN[i];
    if (g2 < 0) continue;
    const int64_t n = g2 >> 1;
#pragma unroll
    for (int c = 0; c < NCOMP; c++) {
      const T a = ACC[c * nb_nodes + i];
      if (g2 & 1) partials[(b * NCOMP + c) * nb_nodes + i] = a;
      else {
        T *dst = v + n + (int64_t)c * comp_stride;
        *dst   = overwrite ? a : *dst + a;
      }
    }
  }
}

}  // namespace mi355
