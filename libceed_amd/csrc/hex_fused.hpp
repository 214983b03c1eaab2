// Fused per-element hex operator kernels for gfx950:
//   gather (E) -> sum-factorised interp / collocated grad (B) -> pointwise D
//   -> B^T -> E^T without atomics: either the element's E-vector block is
//   stored and an ordered gather-sum over the transpose offset map finishes
//   E^T, or elements run in colours that never share a node and add in place.
//
// Thread layout: each element owns a T1 x T1 slab of threads (T1 = Q >= P),
// thread (tx, ty) keeps the z-column of its (x, y) point in registers.  The
// x and y contractions go through LDS slabs (one T1^3 tile per slab), the z
// contraction stays in registers.  Several elements share a workgroup; the 1D
// matrices B (Q x P) and D = G pinv(B) (Q x Q) are staged once per workgroup.
//
// Data layout in HBM (colour order: elements of colour c are contiguous):
//   offsets [elem][P^3]   int32, 2 * node + first_touch (first element, in launch order, reaching the node)
//   q_data  [elem][QC][Q^3]   QC = 1 (mass), 6 (diffusion), 7 (6 + kappa)
//   u, v    L-vectors, component k of node n at n + k*comp_stride
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi355 {

// kQfJit: a user QFunction compiled at run time into the line-layout kernel (operator_jit.hip)
enum { kQfMass = 0, kQfDiffusion = 1, kQfDiffusionKappa = 2, kQfJit = 3 };
// E^T modes: colour-exclusive read-modify-write (ApplyAdd), the same with the
// first-touch overwrite (Apply), hardware atomics (one launch, any order), or
// store of each contribution into an E-vector sorted by L entry (one launch;
// a separate gather sums each L entry's contiguous run of contributions), or
// (diagonal assembly of brick plans) store into an element-major E-vector.
enum { kScatterAdd = 0, kScatterOverwrite = 1, kScatterAtomic = 2, kScatterEvec = 3, kScatterSplit = 4, kScatterElem = 5 };

template <int QF>
struct QfTraits;
template <>
struct QfTraits<kQfMass> {
  static constexpr int kComps = 1, kSlabs = 2;
};
template <>
struct QfTraits<kQfDiffusion> {
  static constexpr int kComps = 6, kSlabs = 3;
};
template <>
struct QfTraits<kQfDiffusionKappa> {
  static constexpr int kComps = 7, kSlabs = 3;
};

template <int P, int Q>
struct HexShape {
  static_assert(Q >= P, "fused hex kernel needs Q >= P (collocated gradient)");
  static constexpr int T1 = Q, NT = Q * Q, SLAB = Q * Q * Q, P3 = P * P * P, Q3 = Q * Q * Q;
};

// LDS entries one workgroup needs: matrices + per-element slabs.
template <int P, int Q, int QF>
__host__ __device__ constexpr int hex_lds_entries(int epb) {
  return Q * P + Q * Q + epb * QfTraits<QF>::kSlabs * Q * Q * Q;
}

// Pointwise Poisson3DApply (jit-source/gallery/ceed-poisson3dapply.h):
// vg_j = sum_i ug_i S_ij with S in Voigt order {xx, yy, zz, yz, xz, xy}.
// The multiply-adds are written out as fmas so that every kernel computing
// it (both thread layouts, with q_data from registers or loaded in place)
// rounds identically, whatever contraction the compiler would choose.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float  fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ void poisson_regs(const T *s, T dx, T dy, T dz, T &vx, T &vy, T &vz) {
  vx = fma_t(dz, s[4], fma_t(dy, s[5], dx * s[0]));
  vy = fma_t(dz, s[3], fma_t(dy, s[1], dx * s[5]));
  vz = fma_t(dz, s[2], fma_t(dy, s[3], dx * s[4]));
}

// The same with the six entries loaded in place (stride between entries);
// q_data is streamed once per apply: non-temporal loads keep it from
// displacing u and v in the caches.
template <typename T>
__device__ __forceinline__ void poisson_point(const T *qd, int64_t stride, T dx, T dy, T dz, T &vx, T &vy, T &vz) {
  T s[6];
#pragma unroll
  for (int k = 0; k < 6; k++) s[k] = __builtin_nontemporal_load(qd + k * stride);
  poisson_regs<T>(s, dx, dy, dz, vx, vy, vz);
}

// Workgroup -> element chunk such that the chunks one XCD runs (workgroups are
// dealt to the 8 XCDs round-robin) are contiguous in the element order.
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t grid) {
  const int64_t x = b % 8, k = b / 8, q = grid / 8, r = grid % 8;
  return x * q + (x < r ? x : r) + k;
}

template <int P, int Q, int NCOMP, int QF, typename T>
__global__ void __launch_bounds__(256) hex_apply_kernel(int64_t elem_begin, int64_t elem_end, int scatter,
                                                         const int32_t *__restrict__ offsets, int32_t comp_stride,
                                                         const T *__restrict__ q_data, const T *__restrict__ mats, const T *__restrict__ u,
                                                         T *__restrict__ v, T *__restrict__ evec, const int32_t *__restrict__ tpos,
                                                         const int32_t *__restrict__ elem_sys,
                                                         const int32_t *__restrict__ sys_active) {
  using S = HexShape<P, Q>;
  constexpr int T1 = S::T1, NT = S::NT, SLAB = S::SLAB, P3 = S::P3, Q3 = S::Q3, QC = QfTraits<QF>::kComps;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *sB = reinterpret_cast<T *>(lds_raw);
  T *sD = sB + Q * P;
  // uniform (z-direction) matrix entries are read from global memory with
  // wave-uniform addresses: scalar loads into SGPRs, no LDS traffic.  The
  // diffusion kernels from P = 8 need more SGPRs than a wave has for them and
  // spill to scratch, so those read the z entries from LDS like x and y.
  constexpr bool kScalarZ = P <= 7 || QF == kQfMass;
  // The column's q_data loads are issued all at once at the start (their
  // latency overlaps the gather and the interpolation) where the registers
  // holding them keep three or more waves per SIMD: mass kernels and diffusion
  // up to Q = 6.  Larger diffusion kernels load each q_z plane where it is used.
  constexpr bool kQdEarly = QF == kQfMass || Q <= 6;
  const T *gB = kScalarZ ? mats : sB, *gD = kScalarZ ? mats + Q * P : sD;

  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  const int tx = t % T1, ty = t / T1;
  T        *S0 = sD + Q * Q + (size_t)le * QfTraits<QF>::kSlabs * SLAB;
  T        *S1 = S0 + SLAB;
  T        *S2 = S1 + SLAB;  // only used by the diffusion variants
  for (int i = threadIdx.x; i < Q * P + Q * Q; i += blockDim.x) sB[i] = mats[i];

  const int64_t e    = elem_begin + xcd_chunk(blockIdx.x, gridDim.x) * epb + le;
  // elements of systems a batched solver has frozen are skipped (elem_sys == nullptr: none)
  const bool live = le < epb && e < elem_end && (!elem_sys || sys_active[elem_sys[e]]);
  if (elem_sys && !__syncthreads_or(live)) return;  // every element of the workgroup is frozen (uniform exit)
  const bool    node = live && tx < P && ty < P;
  auto sidx = [](int z, int y, int x) { return (z * T1 + y) * T1 + x; };

  int off[P];  // 2 * node + first_touch
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) off[z] = offsets[e * P3 + (z * P + ty) * P + tx];
  }
  const T *qd = q_data + (live ? e : 0) * (int64_t)QC * Q3 + (ty * Q + tx);
  // the column's geometric factors, shared by the components (dead lanes read element 0's, unused)
  T sq[Q][QC];
  if constexpr (kQdEarly) {
#pragma unroll
    for (int qz = 0; qz < Q; qz++)
#pragma unroll
      for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
  }
  __syncthreads();

  for (int comp = 0; comp < NCOMP; comp++) {
    const int64_t coff = (int64_t)comp * comp_stride;
    // ---- E: gather the element's nodes of this component into slab S0[z][y][x]
    if (node) {
#pragma unroll
      for (int z = 0; z < P; z++) S0[sidx(z, ty, tx)] = u[(off[z] >> 1) + coff];
    }
    __syncthreads();
    // ---- B, x direction: S1[z][y][qx] = sum_b B[qx][b] S0[z][y][b]   (y, z < P)
    if (ty < P) {
      T acc[P];
#pragma unroll
      for (int z = 0; z < P; z++) acc[z] = 0;
#pragma unroll
      for (int b = 0; b < P; b++) {
        const T bb = sB[tx * P + b];
#pragma unroll
        for (int z = 0; z < P; z++) acc[z] += bb * S0[sidx(z, ty, b)];
      }
#pragma unroll
      for (int z = 0; z < P; z++) S1[sidx(z, ty, tx)] = acc[z];
    }
    __syncthreads();
    // ---- B, y then z direction (z in registers): uq[qz] at point (tx, ty, qz)
    T uq[Q];
    {
      T acc[P];
#pragma unroll
      for (int z = 0; z < P; z++) acc[z] = 0;
#pragma unroll
      for (int b = 0; b < P; b++) {
        const T bb = sB[ty * P + b];
#pragma unroll
        for (int z = 0; z < P; z++) acc[z] += bb * S1[sidx(z, b, tx)];
      }
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += gB[qz * P + b] * acc[b];
        uq[qz] = s;
      }
    }
    T t2[P];  // z-direction B^T result, per (tx, ty) column
    if constexpr (QF == kQfMass) {
      // ---- D: v = u * (w det J)
#pragma unroll
      for (int qz = 0; qz < Q; qz++) uq[qz] *= live ? sq[qz][0] : T(0);
#pragma unroll
      for (int b = 0; b < P; b++) {
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += gB[qz * P + b] * uq[qz];
        t2[b] = s;
      }
    } else {
      // ---- collocated gradient: x, y through LDS, z in registers
#pragma unroll
      for (int qz = 0; qz < Q; qz++) S0[sidx(qz, ty, tx)] = uq[qz];
      __syncthreads();
      T vz[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T dx = 0, dy = 0, dz = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) {
          dx += sD[tx * Q + c] * S0[sidx(qz, ty, c)];
          dy += sD[ty * Q + c] * S0[sidx(qz, c, tx)];
          dz += gD[qz * Q + c] * uq[c];
        }
        T vx = 0, vy = 0, vzz = 0;
        if constexpr (kQdEarly) {
          poisson_regs<T>(sq[qz], dx, dy, dz, vx, vy, vzz);
          if constexpr (QF == kQfDiffusionKappa) {
            vx *= sq[qz][6];
            vy *= sq[qz][6];
            vzz *= sq[qz][6];
          }
          if (!live) vx = vy = vzz = 0;
        } else if (live) {
          const T *q = qd + qz * Q * Q;
          poisson_point<T>(q, Q3, dx, dy, dz, vx, vy, vzz);
          if constexpr (QF == kQfDiffusionKappa) {
            const T k = __builtin_nontemporal_load(q + 6 * Q3);
            vx *= k;
            vy *= k;
            vzz *= k;
          }
        }
        S1[sidx(qz, ty, tx)] = vx;
        S2[sidx(qz, ty, tx)] = vy;
        vz[qz]               = vzz;
      }
      __syncthreads();
      // ---- G^T: sum of the three direction transposes, then B^T in z
      // (x and y sums first, then the z terms: the order the line layout uses)
      T acc[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T sx = 0, sy = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) {
          sx += sD[a * Q + tx] * S1[sidx(qz, ty, a)];
          sy += sD[a * Q + ty] * S2[sidx(qz, a, tx)];
        }
        T s = sx + sy;
#pragma unroll
        for (int a = 0; a < Q; a++) s += gD[a * Q + qz] * vz[a];
        acc[qz] = s;
      }
#pragma unroll
      for (int b = 0; b < P; b++) {
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += gB[qz * P + b] * acc[qz];
        t2[b] = s;
      }
    }
    // ---- B^T, y direction: S1[z][y][qx] = sum_qy B[qy][y] S0[z][qy][qx]   (y, z < P)
#pragma unroll
    for (int b = 0; b < P; b++) S0[sidx(b, ty, tx)] = t2[b];
    __syncthreads();
    if (ty < P) {
      T acc[P];
#pragma unroll
      for (int z = 0; z < P; z++) acc[z] = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) {
        const T bb = sB[a * P + ty];
#pragma unroll
        for (int z = 0; z < P; z++) acc[z] += bb * S0[sidx(z, a, tx)];
      }
#pragma unroll
      for (int z = 0; z < P; z++) S1[sidx(z, ty, tx)] = acc[z];
    }
    __syncthreads();
    // ---- B^T, x direction, then E^T (no atomics unless asked: colour-exclusive
    // read-modify-write, or stores into the E-vector sorted by L entry)
    if (node) {
      T s[P];
#pragma unroll
      for (int z = 0; z < P; z++) {
        T acc = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) acc += sB[a * P + tx] * S1[sidx(z, ty, a)];
        s[z] = acc;
      }
      if (scatter == kScatterEvec) {  // tpos gives each contribution's slot in the sorted E-vector
        int32_t pos[P];
#pragma unroll
        for (int z = 0; z < P; z++) pos[z] = tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + ty) * P + tx];
#pragma unroll
        for (int z = 0; z < P; z++) evec[pos[z]] = s[z];
      } else if (scatter == kScatterAtomic) {  // all elements in one launch
#pragma unroll
        for (int z = 0; z < P; z++) unsafeAtomicAdd(&v[(off[z] >> 1) + coff], s[z]);
      } else if (scatter == kScatterSplit) {  // shared entries to their E-vector slot, the rest onto the zeroed target
#pragma unroll
        for (int z = 0; z < P; z++) {
          if (off[z] & 1) evec[tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + ty) * P + tx]] = s[z];
          else unsafeAtomicAdd(&v[(off[z] >> 1) + coff], s[z]);
        }
      } else {
        // all reads of v before the writes, so the loads are in flight together;
        // Apply's first element to reach an entry (first-touch bit) writes it without reading
        const bool overwrite = scatter == kScatterOverwrite;
        T          old[P];
#pragma unroll
        for (int z = 0; z < P; z++) old[z] = (overwrite && (off[z] & 1)) ? T(0) : v[(off[z] >> 1) + coff];
#pragma unroll
        for (int z = 0; z < P; z++) v[(off[z] >> 1) + coff] = (overwrite && (off[z] & 1)) ? s[z] : old[z] + s[z];
      }
    }
  }
}

}  // namespace mi355
