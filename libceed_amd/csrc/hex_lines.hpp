// Fused per-element hex operator, "line" thread layout.
//
// Same operator as hex_apply_kernel (gather E, sum-factorised interp and
// collocated gradient B, pointwise D, transposes, E^T), with a different
// assignment of work to threads.  The column-slab kernel gives thread (x, y)
// the z-column of its point and contracts x and y through LDS, so every
// output point reads Q values back from LDS (Q-fold redundant reads of the
// same slab row by the Q threads of that row).  Here, for each x- or
// y-contraction, a thread owns one whole 1D line of the slab: it reads the
// line's Q (or P) inputs once, computes all of the line's outputs with the
// 1D matrix entries (wave-uniform: scalar loads, no LDS), and writes them
// back.  Per element and component this reads and writes each slab entry
// about once per contraction instead of Q times, so the LDS traffic, which
// bounds the column-slab kernel at p >= 3, drops about threefold.  The
// z-contractions and the pointwise operator stay in registers on the
// (x, y) columns, as in the column-slab kernel.
//
// Thread t of an element's Q*Q threads plays these roles:
//   column (qx, qy) = (t % Q, t / Q)                     z-contractions, pointwise D
//   node column (x, y) = (t % Q, t / Q), x, y < P      gather E, scatter E^T (coalesced)
//   x-line of nodes (y, z) = (t % P, t / P), t < P*P     B_x, B_x^T
//   y-line (qx, z) = (t % Q, t / Q), t < Q*P              B_y, B_y^T
//   x-line (qy, qz) and y-line (qx, qz) = (t % Q, t / Q)  the collocated derivatives
// Slabs are [z][y][x] with strides (Q*Q, Q, 1), padded at Q = 7, 8 against LDS
// bank conflicts (line_row_stride, line_plane_stride); mass needs two per element,
// diffusion three (the y-derivative transpose waits in registers for a barrier
// and then takes the slab the x-derivative transpose has read: one barrier more
// than four slabs, and a quarter less LDS, so more workgroups fit a CU).
#pragma once

#include "hex_fused.hpp"

namespace mi355 {

// The 1D matrices are read with wave-uniform addresses (scalar loads into
// SGPRs).  Loaded once for the whole kernel they would need ~2(QP + Q^2)
// SGPRs, far more than a wave has, and the compiler would spill them to VGPR
// lanes (a v_writelane/v_readlane pair per use).  Re-deriving the pointer
// through an empty asm at the start of each phase keeps the loads inside the
// phase that uses them, so only that phase's entries are live.
// The pointer is returned in the constant address space (read-only for the
// kernel's lifetime, which the matrices are), so the loads stay scalar.
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ const_ptr<T> phase_ptr(const T *p) {
  asm volatile("" : "+s"(p));
  return (const_ptr<T>)p;
}

// Lanes 2m and 2m+1 of a wave hold the columns (cx, cy) and (cx+1, cy), cx
// even: each loads q_data for both columns, half of the components each, with
// 16-byte loads (half the load instructions), and they swap halves with a DPP
// quad permutation [1, 0, 3, 2] when the values are used.
template <typename T>
using line_v2 = T __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double pair_swap(double x) {
  const long long b  = __double_as_longlong(x);
  const int       lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
  const int       hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int QF>
struct LineSlabs {
  static constexpr int kSlabs = QF == kQfMass ? 2 : 3;
};

// Row and plane strides (entries) of the line kernel's slabs, for entries of
// sb bytes.  At Q = 8 the natural strides (8, 64) are powers of two: the
// entries an x-line, y-line or x-node-line phase touches in one wavefront
// then fall into a few LDS bank groups (SQ_LDS_BANK_CONFLICT 32% of the
// cycles at p=6, 39% with 3 components).  Rows of 9 and planes of 72 (FP64),
// planes of 73 (FP32) spread them (a bank model of the phases' access
// patterns: 2.1x fewer conflict cycles in FP64, 3.8x in FP32).  Measured on
// MI355X, 8M DoFs, ms per apply: FP32 p=6 0.136-0.139 vs 0.148, 3-component
// FP64 p=6 0.648-0.650 vs 0.689-0.691, scalar FP64 p=6 unchanged within run
// to run noise.  At Q = 7 (FP64) planes of 58 model 10% fewer conflict
// cycles: 3-component p=5 0.783-0.784 vs 0.794-0.796 ms, scalar p=5
// unchanged; the same search gave no measurable gain at Q = 4 or 10.
__host__ __device__ constexpr int line_row_stride(int Q, int sb) { return sb == 8 && Q == 8 ? 9 : Q; }
__host__ __device__ constexpr int line_plane_stride(int Q, int sb) { return Q == 8 ? (sb == 8 ? 72 : 73) : Q == 7 && sb == 8 ? 58 : Q * Q; }
__host__ __device__ constexpr int line_slab_entries(int Q, int sb) { return Q * line_plane_stride(Q, sb); }

template <int P, int Q, int QF, typename T>
__host__ __device__ constexpr int hex_lines_lds_entries(int epb) {
  return epb * LineSlabs<QF>::kSlabs * line_slab_entries(Q, (int)sizeof(T));
}

// The column's geometric factors (QC values at each of its Q points) and how
// they are loaded: all loads issued at once and shared by the components, at
// the element start (in flight during E and the interpolation), or for
// scalar FP64 diffusion from Q = 9 just before the pointwise operator, where
// 6Q doubles held through the whole element cost a wave per SIMD (191 -> 137
// VGPRs at Q = 9; measured on MI355X, 8M DoFs: p=7 26.6 -> 27.6 GDoF/s, p=8
// 26.2 -> 27.5; later loads lose at Q <= 8 and in FP32, where the exposed
// latency costs more than the occupancy gains).  FP64 scalar diffusion with
// early loads and Q even loads them in lane pairs with 16-byte loads
// (interleaved A/B on MI355X, 8M DoFs: p=4 +0.7% with colours, +3.7% with the
// map, p=6 +0-4% with the split; with the late loads at Q = 10 the swap in the
// pointwise loop cost 15%; FP32 pairs of 8-byte loads were 3-5% slower).
template <int Q, int NCOMP, int QF, typename T>
struct LineQData {
  static constexpr int  QC     = QfTraits<QF>::kComps, Q3 = Q * Q * Q, QH = QC / 2;
  static constexpr bool kEarly = QF == kQfMass || NCOMP > 1 || Q < 9 || sizeof(T) < 8;
  static constexpr bool kPair  = kEarly && sizeof(T) == 8 && QF == kQfDiffusion && NCOMP == 1 && Q % 2 == 0;
  T                     sq[kPair ? 1 : Q][QC];
  line_v2<T>            sp[kPair ? Q : 1][kPair ? QH : 1];

  // qe: the element's block [QC][Q^3]; (cx, cy) the column
  __device__ __forceinline__ void load(const T *qe, int cx, int cy) {
    if constexpr (kPair) {  // 16-byte aligned: Q even, so (cy * Q + cx - odd) and every stride are even
      const T *pb = qe + (cy * Q + (cx & ~1)) + ((cx & 1) ? QH : 0) * Q3;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
#pragma unroll
        for (int kk = 0; kk < QH; kk++) sp[qz][kk] = __builtin_nontemporal_load((const line_v2<T> *)(pb + kk * Q3 + qz * Q * Q));
    } else {
      const T *qd = qe + (cy * Q + cx);
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
#pragma unroll
        for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
    }
  }
  // the QC values of point qz of the column (with the pairs: this column's
  // half from the own load, the other half from the partner lane)
  __device__ __forceinline__ void point(int qz, int cx, T (&s)[QC]) const {
    if constexpr (kPair) {
      const bool odd = cx & 1;
#pragma unroll
      for (int kk = 0; kk < QH; kk++) {
        const line_v2<T> d = sp[qz][kk];
        const T       keep = odd ? d.y : d.x, got = pair_swap(odd ? d.x : d.y);
        s[kk]              = odd ? got : keep;
        s[QH + kk]         = odd ? keep : got;
      }
    } else {
#pragma unroll
      for (int k = 0; k < QC; k++) s[k] = sq[qz][k];
    }
  }
};

// Synchronisation of one element's threads between the phases of the line
// pipeline: the whole workgroup, or the wavefront alone when an element never
// spans two wavefronts (a wavefront's LDS operations are performed in the
// order it issues them; the fences keep the compiler from moving them).
struct SyncBlock {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
};
struct SyncWave {
  __device__ __forceinline__ void operator()() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// One element and component through B, D and B^T in the line layout.  On
// entry S0 holds the element's nodal values at (z, y, x), x, y, z < P, and
// the element's threads are synchronised; on return S0 holds the nodal
// results there, synchronised again.  t is the thread's index among the
// element's Q*Q threads; threads with active false (lanes of a wavefront
// that hold no element) only take part in the synchronisations.  qd holds
// the column's geometric factors (loaded here first when they are late,
// from qe); live false zeroes the pointwise output.
template <int P, int Q, int NCOMP, int QF, typename T, typename Sync>
__device__ __forceinline__ void line_element(T *S0, T *S1, T *S2, int t, bool active, bool live, const T *__restrict__ mats,
                                             LineQData<Q, NCOMP, QF, T> &qd, const T *qe, Sync sync) {
  constexpr int QC = QfTraits<QF>::kComps;
  // from Q = 7 one phase's matrix entries no longer fit the SGPRs: load them
  // per output row (or column) instead of per phase
  constexpr bool kRowWise = Q >= 7;
  // B (Q x P) then D (Q x Q), row-major, at mats
  auto       at = [](int z, int y, int x) { return (z * Q + y) * Q + x; };
  const int  cx = t % Q, cy = t / Q;  // column
  const int  ny = t % P, nz = t / P;  // x-line of nodes
  const bool xline = active && t < P * P;
  // ---- B_x: x-line (ny, nz) of nodes, Q interpolated values into S1
  if (xline) {
    const auto gB = phase_ptr(mats);
    T          in[P];
#pragma unroll
    for (int x = 0; x < P; x++) in[x] = S0[at(nz, ny, x)];
#pragma unroll
    for (int qx = 0; qx < Q; qx++) {
      const auto m = kRowWise ? phase_ptr(mats + qx * P) : gB + qx * P;
      T          s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * in[b];
      S1[at(nz, ny, qx)] = s;
    }
  }
  sync();
  // ---- B_y: y-line (qx, z), z < P
  if (active && t < Q * P) {
    const auto gB = phase_ptr(mats);
    const int  qx = t % Q, z = t / Q;
    T          in[P];
#pragma unroll
    for (int b = 0; b < P; b++) in[b] = S1[at(z, b, qx)];
#pragma unroll
    for (int qy = 0; qy < Q; qy++) {
      const auto m = kRowWise ? phase_ptr(mats + qy * P) : gB + qy * P;
      T          s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * in[b];
      S0[at(z, qy, qx)] = s;
    }
  }
  sync();
  // ---- B_z on the column (cx, cy): uq[qz]
  T uq[Q];
  if (active) {
    const auto gB = phase_ptr(mats);
    T          col[P];
#pragma unroll
    for (int b = 0; b < P; b++) col[b] = S0[at(b, cy, cx)];
#pragma unroll
    for (int qz = 0; qz < Q; qz++) {
      const auto m = kRowWise ? phase_ptr(mats + qz * P) : gB + qz * P;
      T          s = 0;
#pragma unroll
      for (int b = 0; b < P; b++) s += m[b] * col[b];
      uq[qz] = s;
    }
  }
  if constexpr (QF == kQfMass) {
    // ---- D, then B_z^T back into the column's own S0 entries (read only by this thread)
    if (active) {
      const auto gB = phase_ptr(mats);
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T s[QC];
        qd.point(qz, cx, s);
        uq[qz] *= live ? s[0] : T(0);
      }
#pragma unroll
      for (int b = 0; b < P; b++) {
        const auto m = kRowWise ? phase_ptr(mats + b) : gB + b;
        T          s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += m[qz * P] * uq[qz];
        S0[at(b, cy, cx)] = s;
      }
    }
  } else {
    // ---- collocated derivatives: z on the column, x and y by lines through S1
    T dz[Q];
    if (active) {
      const auto gD = phase_ptr(mats + Q * P);
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qz * Q) : gD + qz * Q;
        T          s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * uq[c];
        dz[qz]             = s;
        S1[at(qz, cy, cx)] = uq[qz];
      }
    }
    sync();
    if (active) {
      const auto gD = phase_ptr(mats + Q * P);
      const int  a = t % Q, qz = t / Q;  // x-line (qy = a, qz) and y-line (qx = a, qz)
      T          in[Q];
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S1[at(qz, a, c)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qx * Q) : gD + qx * Q;
        T          s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        S2[at(qz, a, qx)] = s;
      }
#pragma unroll
      for (int c = 0; c < Q; c++) in[c] = S1[at(qz, c, a)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qy * Q) : gD + qy * Q;
        T          s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * in[c];
        S0[at(qz, qy, a)] = s;
      }
    }
    sync();
    // ---- pointwise (Poisson3DApply, times kappa for the kappa-weighted operator) on the column
    if constexpr (!LineQData<Q, NCOMP, QF, T>::kEarly) {
      if (active) qd.load(qe, cx, cy);
    }
    T vz[Q];
    if (active) {
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T vx, vy, vzz;
        T s[QC];
        qd.point(qz, cx, s);
        poisson_regs<T>(s, S2[at(qz, cy, cx)], S0[at(qz, cy, cx)], dz[qz], vx, vy, vzz);
        if constexpr (QF == kQfDiffusionKappa) {
          vx *= s[6];
          vy *= s[6];
          vzz *= s[6];
        }
        if (!live) vx = vy = vzz = 0;
        S2[at(qz, cy, cx)] = vx;
        S0[at(qz, cy, cx)] = vy;
        vz[qz]             = vzz;
      }
    }
    sync();
    // ---- D_x^T by x-lines into S1; D_y^T by y-lines, held in registers
    // until every x-line has read S2, then written over it
    {
      const int a = t % Q, qz = t / Q;
      T         ty[Q];
      if (active) {
        {
          const auto gD = phase_ptr(mats + Q * P);
          T          in[Q];
#pragma unroll
          for (int c = 0; c < Q; c++) in[c] = S2[at(qz, a, c)];
#pragma unroll
          for (int qx = 0; qx < Q; qx++) {
            const auto m = kRowWise ? phase_ptr(mats + Q * P + qx) : gD + qx;
            T          s = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
            S1[at(qz, a, qx)] = s;
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
            T          s = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
            ty[qy] = s;
          }
        }
      }
      sync();
      if (active) {
#pragma unroll
        for (int qy = 0; qy < Q; qy++) S2[at(qz, qy, a)] = ty[qy];
      }
    }
    sync();
    // ---- sum of the three transposes on the column, then B_z^T into S0
    if (active) {
      const auto gDz = phase_ptr(mats + Q * P);
      T          w[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qz) : gDz + qz;
        T          s = S1[at(qz, cy, cx)] + S2[at(qz, cy, cx)];
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c * Q] * vz[c];
        w[qz] = s;
      }
      const auto gBz = phase_ptr(mats);
#pragma unroll
      for (int b = 0; b < P; b++) {
        const auto m = kRowWise ? phase_ptr(mats + b) : gBz + b;
        T          s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += m[qz * P] * w[qz];
        S0[at(b, cy, cx)] = s;
      }
    }
  }
  sync();
  // ---- B_y^T: y-line (qx, z), z < P, into S1
  if (active && t < Q * P) {
    const auto gB = phase_ptr(mats);
    const int  qx = t % Q, z = t / Q;
    T          in[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) in[a] = S0[at(z, a, qx)];
#pragma unroll
    for (int y = 0; y < P; y++) {
      const auto m = kRowWise ? phase_ptr(mats + y) : gB + y;
      T          s = 0;
#pragma unroll
      for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
      S1[at(z, y, qx)] = s;
    }
  }
  sync();
  // ---- B_x^T: x-line (ny, nz) of nodes, into S0 (last read by B_y^T)
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
  sync();
}

// The line kernel keeps the pipeline written out in its own body rather than
// calling line_element: the same code through the shared function compiles
// to a different register allocation here (P=7, Q=8 diffusion: 186 against
// 162 VGPRs, a wave per SIMD fewer, 20% slower on MI355X at 8M DoFs; P=5,
// Q=6: 133 against 127).  Both compute every value with the same operations
// in the same order (tests/test_gpu_brick.py::test_brick_single_brick_bitwise).
// kSeg: the workgroup runs one segment of elements chained along x
// (seg_start), summing the face between consecutive elements in LDS before
// E^T; a template parameter so that the other modes keep their registers.
// Waves per SIMD the register allocation must leave room for (1: no bound).
// Only where the bound costs no spills and measured faster: FP64 3-component
// diffusion at P = 4, Q = 5 fits 4 waves in 115 VGPRs instead of 3 in 129
// (MI355X, 8M nodes: 0.779-0.789 vs 0.826-0.830 ms per apply).  Measured and
// not kept: 3-component FP32 at P = 6, Q = 7 (5 waves, no spills) 4% slower;
// FP64 3-component at P = 5, Q = 6 (4 waves, 8 VGPRs spilled) 7% slower.
template <int P, int Q, int NCOMP, int QF, typename T>
constexpr int lines_min_waves() {
  return sizeof(T) == 8 && QF == kQfDiffusion && NCOMP == 3 && P == 4 && Q == 5 ? 4 : 1;
}
template <int P, int Q, int NCOMP, int QF, typename T, bool kSeg = false>
__global__ void __launch_bounds__(256, (lines_min_waves<P, Q, NCOMP, QF, T>())) hex_apply_lines_kernel(int64_t elem_begin, int64_t elem_end, int scatter,
                                                               const int32_t *__restrict__ offsets, int32_t comp_stride,
                                                               const T *__restrict__ q_data, const T *__restrict__ mats,
                                                               const T *__restrict__ u, T *__restrict__ v, T *__restrict__ evec,
                                                               const int32_t *__restrict__ tpos,
                                                               const int32_t *__restrict__ elem_sys, const int32_t *__restrict__ sys_active,
                                                               const int32_t *__restrict__ seg_start) {
  constexpr int NT = Q * Q, P3 = P * P * P, Q3 = Q * Q * Q, QC = QfTraits<QF>::kComps;
  constexpr int SR = line_row_stride(Q, sizeof(T)), SP = line_plane_stride(Q, sizeof(T)), SLAB = Q * SP;
  static_assert(SR >= Q && SP >= Q * SR, "slab rows and planes must not overlap");
  constexpr int NS = LineSlabs<QF>::kSlabs;
  // from Q = 7 one phase's matrix entries no longer fit the SGPRs: load them
  // per output row (or column) instead of per phase
  constexpr bool kRowWise = Q >= 7;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  // B (Q x P) then D (Q x Q), row-major, at mats

  const int epb = blockDim.x / NT;
  const int le = threadIdx.x / NT, t = threadIdx.x - le * NT;
  T        *S0 = reinterpret_cast<T *>(lds_raw) + (size_t)le * NS * SLAB;
  T        *S1 = S0 + SLAB;
  T        *S2 = S1 + SLAB;  // diffusion only
  auto      at = [](int z, int y, int x) { return z * SP + y * SR + x; };

  // the workgroup's elements: epb consecutive packed elements, or (kSeg)
  // segment elem_begin + chunk, whose consecutive elements share x-faces
  const int64_t chunk = xcd_chunk(blockIdx.x, gridDim.x);
  const int32_t sbeg  = kSeg ? seg_start[elem_begin + chunk] : 0;
  const int32_t slen  = kSeg ? seg_start[elem_begin + chunk + 1] - sbeg : 0;
  const int64_t e     = kSeg ? (int64_t)sbeg + le : elem_begin + chunk * epb + le;
  const bool    live  = le < epb && (kSeg ? le < slen : e < elem_end) && (!elem_sys || sys_active[elem_sys[e]]);
  if (elem_sys && !__syncthreads_or(live)) return;  // every element of the workgroup is frozen (uniform exit)

  const int  cx = t % Q, cy = t / Q;  // column
  const int  ny = t % P, nz = t / P;  // x-line of nodes
  const bool xline = t < P * P;
  // E and E^T run on the node columns (cx, cy < P), like the column layout, so
  // that a wave's loads and stores of u and v are coalesced along x
  const bool node = live && cx < P && cy < P;
  int        off[P];  // 2 * node + first_touch, nodes (cx, cy, z)
  if (node) {
#pragma unroll
    for (int z = 0; z < P; z++) off[z] = offsets[e * P3 + (z * P + cy) * P + cx];
  }
  // the column's geometric factors, all loads issued at once and shared by the
  // components: at the element start (in flight during E and the
  // interpolation), or for scalar FP64 diffusion from Q = 9 just before the
  // pointwise operator, where 6Q doubles held through the whole element cost
  // a wave per SIMD (191 -> 137 VGPRs at Q = 9; measured on MI355X, 8M DoFs:
  // p=7 26.6 -> 27.6 GDoF/s, p=8 26.2 -> 27.5; later loads lose at Q <= 8
  // and in FP32, where the exposed latency costs more than the occupancy gains)
  constexpr bool kQdEarly = QF == kQfMass || NCOMP > 1 || Q < 9 || sizeof(T) < 8;
  // paired 16-byte loads with the early loads of FP64 scalar diffusion, Q even
  // (interleaved A/B on MI355X, 8M DoFs: p=4 +0.7% with colours, +3.7% with
  // the map, p=6 +0-4% with the split; with the late loads at Q = 10 the swap
  // in the pointwise loop cost 15%; FP32 pairs of 8-byte loads were 3-5% slower)
  constexpr bool kPairQd = kQdEarly && sizeof(T) == 8 && QF == kQfDiffusion && NCOMP == 1 && Q % 2 == 0;
  constexpr int  QH      = QC / 2;
  const T       *qd       = q_data + (live ? e : 0) * (int64_t)QC * Q3 + (cy * Q + cx);
  T              sq[kPairQd ? 1 : Q][QC];
  line_v2<T>     sp[kPairQd ? Q : 1][kPairQd ? QH : 1];
  auto           load_q = [&]() {
    if constexpr (kPairQd) {  // 16-byte aligned: Q even, so (cy * Q + cx - odd) and every stride are even
      const T *pb = q_data + (live ? e : 0) * (int64_t)QC * Q3 + (cy * Q + (cx & ~1)) + ((cx & 1) ? QH : 0) * Q3;
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
#pragma unroll
        for (int kk = 0; kk < QH; kk++) sp[qz][kk] = __builtin_nontemporal_load((const line_v2<T> *)(pb + kk * Q3 + qz * Q * Q));
    } else {
#pragma unroll
      for (int qz = 0; qz < Q; qz++)
#pragma unroll
        for (int k = 0; k < QC; k++) sq[qz][k] = __builtin_nontemporal_load(&qd[k * Q3 + qz * Q * Q]);
    }
  };
  if constexpr (kQdEarly) load_q();

  for (int comp = 0; comp < NCOMP; comp++) {
    const int64_t coff = (int64_t)comp * comp_stride;
    if (comp > 0) __syncthreads();  // the previous component's last phase still reads S0
    // ---- E: the node column gathers its P nodes into S0
    if (node) {
#pragma unroll
      for (int z = 0; z < P; z++) S0[at(z, cy, cx)] = u[(off[z] >> 1) + coff];
    }
    __syncthreads();
    // ---- B_x: x-line (ny, nz) of nodes, Q interpolated values into S1
    if (xline) {
      const auto gB = phase_ptr(mats);
      T        in[P];
#pragma unroll
      for (int x = 0; x < P; x++) in[x] = S0[at(nz, ny, x)];
#pragma unroll
      for (int qx = 0; qx < Q; qx++) {
        const auto m = kRowWise ? phase_ptr(mats + qx * P) : gB + qx * P;
        T s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * in[b];
        S1[at(nz, ny, qx)] = s;
      }
    }
    __syncthreads();
    // ---- B_y: y-line (qx, z), z < P
    if (t < Q * P) {
      const auto gB = phase_ptr(mats);
      const int qx = t % Q, z = t / Q;
      T         in[P];
#pragma unroll
      for (int b = 0; b < P; b++) in[b] = S1[at(z, b, qx)];
#pragma unroll
      for (int qy = 0; qy < Q; qy++) {
        const auto m = kRowWise ? phase_ptr(mats + qy * P) : gB + qy * P;
        T s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * in[b];
        S0[at(z, qy, qx)] = s;
      }
    }
    __syncthreads();
    // ---- B_z on the column (cx, cy): uq[qz]
    T uq[Q];
    {
      const auto gB = phase_ptr(mats);
      T        col[P];
#pragma unroll
      for (int b = 0; b < P; b++) col[b] = S0[at(b, cy, cx)];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + qz * P) : gB + qz * P;
        T s = 0;
#pragma unroll
        for (int b = 0; b < P; b++) s += m[b] * col[b];
        uq[qz] = s;
      }
    }
    if constexpr (QF == kQfMass) {
      // ---- D, then B_z^T back into the column's own S0 entries (read only by this thread)
      const auto gB = phase_ptr(mats);
#pragma unroll
      for (int qz = 0; qz < Q; qz++) uq[qz] *= live ? sq[qz][0] : T(0);
#pragma unroll
      for (int b = 0; b < P; b++) {
        const auto m = kRowWise ? phase_ptr(mats + b) : gB + b;
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += m[qz * P] * uq[qz];
        S0[at(b, cy, cx)] = s;
      }
    } else {
      // ---- collocated derivatives: z on the column, x and y by lines through S1
      const auto gD = phase_ptr(mats + Q * P);
      T        dz[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qz * Q) : gD + qz * Q;
        T s = 0;
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c] * uq[c];
        dz[qz] = s;
        S1[at(qz, cy, cx)] = uq[qz];
      }
      __syncthreads();
      {
        const auto gD = phase_ptr(mats + Q * P);
        const int a = t % Q, qz = t / Q;  // x-line (qy = a, qz) and y-line (qx = a, qz)
        T         in[Q];
#pragma unroll
        for (int c = 0; c < Q; c++) in[c] = S1[at(qz, a, c)];
#pragma unroll
        for (int qx = 0; qx < Q; qx++) {
          const auto m = kRowWise ? phase_ptr(mats + Q * P + qx * Q) : gD + qx * Q;
          T s = 0;
#pragma unroll
          for (int c = 0; c < Q; c++) s += m[c] * in[c];
          S2[at(qz, a, qx)] = s;
        }
#pragma unroll
        for (int c = 0; c < Q; c++) in[c] = S1[at(qz, c, a)];
#pragma unroll
        for (int qy = 0; qy < Q; qy++) {
          const auto m = kRowWise ? phase_ptr(mats + Q * P + qy * Q) : gD + qy * Q;
          T s = 0;
#pragma unroll
          for (int c = 0; c < Q; c++) s += m[c] * in[c];
          S0[at(qz, qy, a)] = s;
        }
      }
      __syncthreads();
      // ---- pointwise (Poisson3DApply, times kappa for the kappa-weighted operator) on the column
      // From Q = 10 the column's z-derivatives wait in S1 (free until D_x^T
      // below; only this thread's entries) rather than in registers beside
      // the late q_data loads (interleaved A/B on MI355X, 8M DoFs: p=8 0.283
      // vs 0.293 ms; at Q = 9, p=7, 0.280 vs 0.274, so not there).
      constexpr bool kDzLds = !kQdEarly && Q >= 10;
      if constexpr (kDzLds) {
#pragma unroll
        for (int qz = 0; qz < Q; qz++) S1[at(qz, cy, cx)] = dz[qz];
      }
      if constexpr (!kQdEarly) load_q();
      T vz[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        T vx, vy, vzz;
        if constexpr (kPairQd) {  // this column's half from the own load, the other half from the partner lane
          const bool odd = cx & 1;
          T          s6[QC];
#pragma unroll
          for (int kk = 0; kk < QH; kk++) {
            const line_v2<T> d = sp[qz][kk];
            const T       keep = odd ? d.y : d.x, got = pair_swap(odd ? d.x : d.y);
            s6[kk]             = odd ? got : keep;
            s6[QH + kk]        = odd ? keep : got;
          }
          poisson_regs<T>(s6, S2[at(qz, cy, cx)], S0[at(qz, cy, cx)], dz[qz], vx, vy, vzz);
        } else {
          poisson_regs<T>(sq[qz], S2[at(qz, cy, cx)], S0[at(qz, cy, cx)], kDzLds ? S1[at(qz, cy, cx)] : dz[qz], vx, vy, vzz);
        }
        if constexpr (QF == kQfDiffusionKappa) {
          vx *= sq[qz][6];
          vy *= sq[qz][6];
          vzz *= sq[qz][6];
        }
        if (!live) vx = vy = vzz = 0;
        S2[at(qz, cy, cx)] = vx;
        S0[at(qz, cy, cx)] = vy;
        vz[qz]             = vzz;
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
            T s = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
            S1[at(qz, a, qx)] = s;
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
            T s = 0;
#pragma unroll
            for (int c = 0; c < Q; c++) s += m[c * Q] * in[c];
            ty[qy] = s;
          }
        }
        __syncthreads();
#pragma unroll
        for (int qy = 0; qy < Q; qy++) S2[at(qz, qy, a)] = ty[qy];
      }
      __syncthreads();
      // ---- sum of the three transposes on the column, then B_z^T into S0
      const auto gDz = phase_ptr(mats + Q * P);
      T        w[Q];
#pragma unroll
      for (int qz = 0; qz < Q; qz++) {
        const auto m = kRowWise ? phase_ptr(mats + Q * P + qz) : gDz + qz;
        T s = S1[at(qz, cy, cx)] + S2[at(qz, cy, cx)];
#pragma unroll
        for (int c = 0; c < Q; c++) s += m[c * Q] * vz[c];
        w[qz] = s;
      }
      const auto gBz = phase_ptr(mats);
#pragma unroll
      for (int b = 0; b < P; b++) {
        const auto m = kRowWise ? phase_ptr(mats + b) : gBz + b;
        T s = 0;
#pragma unroll
        for (int qz = 0; qz < Q; qz++) s += m[qz * P] * w[qz];
        S0[at(b, cy, cx)] = s;
      }
    }
    __syncthreads();
    // ---- B_y^T: y-line (qx, z), z < P, into S1
    if (t < Q * P) {
      const auto gB = phase_ptr(mats);
      const int qx = t % Q, z = t / Q;
      T         in[Q];
#pragma unroll
      for (int a = 0; a < Q; a++) in[a] = S0[at(z, a, qx)];
#pragma unroll
      for (int y = 0; y < P; y++) {
        const auto m = kRowWise ? phase_ptr(mats + y) : gB + y;
        T s = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) s += m[a * P] * in[a];
        S1[at(z, y, qx)] = s;
      }
    }
    __syncthreads();
    // ---- B_x^T: x-line (ny, nz) of nodes, into S0 (last read by B_y^T)
    if (xline) {
      const auto gB = phase_ptr(mats);
      T        in[Q];
#pragma unroll
      for (int a = 0; a < Q; a++) in[a] = S1[at(nz, ny, a)];
#pragma unroll
      for (int x = 0; x < P; x++) {
        const auto m = kRowWise ? phase_ptr(mats + x) : gB + x;
        T acc = 0;
#pragma unroll
        for (int a = 0; a < Q; a++) acc += m[a * P] * in[a];
        S0[at(nz, ny, x)] = acc;
      }
    }
    __syncthreads();
    // ---- E^T from the node columns
    if (node) {
      T s[P];
#pragma unroll
      for (int z = 0; z < P; z++) s[z] = S0[at(z, cy, cx)];
      if (!kSeg && scatter == kScatterEvec) {  // tpos gives each contribution's slot in the E-vector sorted by L entry
        int32_t pos[P];
#pragma unroll
        for (int z = 0; z < P; z++) pos[z] = tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + cy) * P + cx];
#pragma unroll
        for (int z = 0; z < P; z++) evec[pos[z]] = s[z];
      } else if (!kSeg && scatter == kScatterAtomic) {
#pragma unroll
        for (int z = 0; z < P; z++) unsafeAtomicAdd(&v[(off[z] >> 1) + coff], s[z]);
      } else if (!kSeg && scatter == kScatterSplit) {  // shared entries to their E-vector slot, the rest onto the zeroed target
#pragma unroll
        for (int z = 0; z < P; z++) {
          if (off[z] & 1) evec[tpos[(e * NCOMP + comp) * (int64_t)P3 + (z * P + cy) * P + cx]] = s[z];
          else unsafeAtomicAdd(&v[(off[z] >> 1) + coff], s[z]);
        }
      } else {
        // segments: the face x = P-1 also holds the next element's x = 0 face
        // (its S0 is final after the barrier above), added after the own value;
        // the next element's threads of that face write nothing
        bool write = true;
        if constexpr (kSeg) {
          if (cx == P - 1 && le + 1 < slen) {
            const T *Sn = S0 + NS * SLAB;
#pragma unroll
            for (int z = 0; z < P; z++) s[z] += Sn[at(z, cy, 0)];
          }
          write = cx > 0 || le == 0;
        }
        if (write) {
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
}

}  // namespace mi355
