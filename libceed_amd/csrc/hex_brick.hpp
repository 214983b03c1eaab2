// Fused per-element hex operator, "brick" E^T.
//
// The plan groups the elements into bricks of up to bx * by * bz face-sharing
// elements (2 x 2 x 2 on a structured mesh).  One workgroup owns one brick
// and runs all of its elements at once, each through the line-layout pipeline
// of hex_apply_lines_kernel (line_element: interp and collocated gradient B,
// pointwise D, transposes), each in its own LDS slabs.  Instead of a global
// E^T per element, every element leaves its P^3 results in its slab, and
// after one workgroup barrier each node of the brick sums the results of the
// brick's elements that contain it, in slot order:
//   * nodes whose every element lies in the brick ("complete") are written to
//     v once (Apply) or added to it (ApplyAdd): no atomics, no re-reads;
//   * the other nodes (on the brick's surface) store their partial sum into a
//     per-brick slot, and brick_gather_kernel adds the partials of each such
//     node in ascending brick order.
// Every sum has a fixed order (slot order inside a brick, brick order across
// bricks), so the result is bitwise reproducible; the order differs from the
// reference's ascending (elem, comp, node) order only by rounding.
//
// The brick's node list (2 * L-node + incomplete bit, or -1 where no element
// of the brick lies) is staged in LDS once; an element's gather reads its
// node indices from it by the brick-local position, so the only global loads
// of the gather are u itself (no per-element offsets are read).
//
// Threads: an element has Q*Q threads.  When that fits a wavefront (Q <= 8),
// each wavefront holds 64 / (Q*Q) whole elements and the element's phases
// synchronise the wavefront only, so the elements of a brick run
// independently between the brick's first and last barrier; from Q = 9 the
// elements are packed back to back and every phase is a workgroup barrier.
#pragma once

#include <type_traits>

#include "hex_lines.hpp"

namespace mi355 {

template <int P, int Q>
struct BrickThreads {
  static constexpr int  NT    = Q * Q;
  static constexpr bool kWave = NT <= 64;                // one-wavefront elements
  static constexpr int  kEpw  = kWave ? 64 / NT : 1;     // elements per wavefront
  // threads of the workgroup of a brick of ns slots
  static constexpr int threads(int ns) { return kWave ? 64 * ((ns + kEpw - 1) / kEpw) : ns * NT; }
};

// LDS bytes of one brick workgroup: ns elements' line-layout slabs and the node list.
template <int P, int Q, int QF, typename T>
__host__ __device__ constexpr size_t hex_brick_lds_bytes(int ns, int nb_nodes) {
  return sizeof(T) * (size_t)ns * LineSlabs<QF>::kSlabs * Q * Q * Q + sizeof(int32_t) * (size_t)nb_nodes;
}

template <int P, int Q, int NCOMP, int QF, typename T>
__global__ void __launch_bounds__(512) hex_apply_brick_kernel(int bx, int by, int bz, int nb_nodes, const int32_t *__restrict__ brick_start,
                                                               const int32_t *__restrict__ brick_mask, const int32_t *__restrict__ bnode,
                                                               int32_t comp_stride, const T *__restrict__ q_data,
                                                               const T *__restrict__ mats, const T *__restrict__ u, T *__restrict__ v,
                                                               T *__restrict__ partials, int overwrite) {
  using BT                = BrickThreads<P, Q>;
  constexpr int NT        = Q * Q, SLAB = Q * Q * Q, Q3 = Q * Q * Q, QC = QfTraits<QF>::kComps;
  constexpr int NS        = LineSlabs<QF>::kSlabs, p = P - 1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  auto          at = [](int z, int y, int x) { return (z * Q + y) * Q + x; };

  const int ns  = bx * by * bz;
  const int tid = threadIdx.x;
  // slot (element of the brick) and index among the element's threads
  int  slot, t;
  bool active;  // false for wavefront lanes that hold no slot of this brick
  if constexpr (BT::kWave) {
    const int lane = tid & 63, k = lane / NT;
    slot           = (tid >> 6) * BT::kEpw + k;
    t              = lane - k * NT;
    active         = k < BT::kEpw && slot < ns;
  } else {
    slot   = tid / NT;
    t      = tid - slot * NT;
    active = true;
  }
  T       *slabs = reinterpret_cast<T *>(lds_raw);
  int32_t *sN    = reinterpret_cast<int32_t *>(slabs + (size_t)ns * NS * SLAB);  // [nb_nodes]
  T       *S0    = slabs + (size_t)(active ? slot : 0) * NS * SLAB;
  T       *S1    = S0 + SLAB;
  T       *S2    = S1 + SLAB;  // diffusion only

  const int64_t b = xcd_chunk(blockIdx.x, gridDim.x);  // grid = number of bricks
  for (int i = tid; i < nb_nodes; i += blockDim.x) sN[i] = bnode[b * nb_nodes + i];
  const int     mask = brick_mask[b];
  const bool    live = active && ((mask >> slot) & 1);
  const int64_t e    = brick_start[b] + __builtin_popcount(mask & ((1u << (active ? slot : 0)) - 1));
  const int     cx = t % Q, cy = t / Q;  // column; node column when both < P
  const bool    node = live && cx < P && cy < P;
  const int     nbx = bx * p + 1, nby = by * p + 1, nbxy = nbx * nby;
  const int     sx = slot % bx, sy = (slot / bx) % by, sz = slot / (bx * by);
  // brick-local index of node (cx, cy, 0) of the element; + z * nbxy for node z
  const int l0 = (sz * p) * nbxy + (sy * p + cy) * nbx + sx * p + cx;

  LineQData<Q, NCOMP, QF, T> qd;
  const T                   *qe = q_data + (live ? e : 0) * (int64_t)QC * Q3;
  if constexpr (LineQData<Q, NCOMP, QF, T>::kEarly) qd.load(qe, cx, cy);
  __syncthreads();  // node list

  using Sync = typename std::conditional<BT::kWave, SyncWave, SyncBlock>::type;
  for (int comp = 0; comp < NCOMP; comp++) {
    const int64_t coff = (int64_t)comp * comp_stride;
    if (comp > 0) __syncthreads();  // the previous component's node sums still read the slabs
    // ---- E: the node column gathers its P nodes into S0
    if (node) {
      T g[P];
#pragma unroll
      for (int z = 0; z < P; z++) g[z] = u[(sN[l0 + z * nbxy] >> 1) + coff];
#pragma unroll
      for (int z = 0; z < P; z++) S0[at(z, cy, cx)] = g[z];
    }
    Sync{}();
    line_element<P, Q, NCOMP, QF, T>(S0, S1, S2, t, active, live, mats, qd, qe, Sync{});
    __syncthreads();  // every element's results in its S0
    // ---- the brick's nodes: the sum over the brick's elements holding the node, in slot order;
    // complete nodes to v, the others to their partial slot
    for (int i = tid; i < nb_nodes; i += blockDim.x) {
      const int32_t g2 = sN[i];
      if (g2 < 0) continue;
      const int bi = i % nbx, bj = (i / nbx) % nby, bk = i / nbxy;
      // slots holding position (bi, bj, bk): one or two along each direction
      const int sx1 = bi / p < bx ? bi / p : bx - 1, sx0 = (bi % p == 0 && bi > 0) ? bi / p - 1 : sx1;
      const int sy1 = bj / p < by ? bj / p : by - 1, sy0 = (bj % p == 0 && bj > 0) ? bj / p - 1 : sy1;
      const int sz1 = bk / p < bz ? bk / p : bz - 1, sz0 = (bk % p == 0 && bk > 0) ? bk / p - 1 : sz1;
      T         acc = 0;
      for (int kz = sz0; kz <= sz1; kz++)
        for (int ky = sy0; ky <= sy1; ky++)
          for (int kx = sx0; kx <= sx1; kx++) {
            const int s = kx + bx * (ky + by * kz);
            if ((mask >> s) & 1) acc += slabs[(size_t)s * NS * SLAB + at(bk - kz * p, bj - ky * p, bi - kx * p)];
          }
      if (g2 & 1) partials[(b * NCOMP + comp) * nb_nodes + i] = acc;
      else {
        T *dst = v + (g2 >> 1) + coff;
        *dst   = overwrite ? acc : *dst + acc;
      }
    }
  }
}

}  // namespace mi355
