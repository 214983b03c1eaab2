// Operator plan shared by the plan code (operator_hex.hip) and the kernel
// instantiation units (hex_apply_f64.hip, hex_apply_f32.hip).
#pragma once

#include <stdint.h>

#include <vector>

#include "hex_brick.hpp"
#include "hex_fused.hpp"
#include "hex_lines.hpp"
#include "hex_points.hpp"
#include "mi355_common.h"

struct CeedMi355Operator_private;

namespace mi355 {

using ApplyFn = int (*)(const CeedMi355Operator_private *op, const void *u, void *v, int overwrite, hipStream_t s);
// v (=|+=) E^T evec by the plan's transpose map (operator_hex.hip)
int transpose_gather(const CeedMi355Operator_private *op, void *v, int overwrite, hipStream_t s);
// The split mode's second step: Apply (overwrite) writes the shared entries of
// v; ApplyAdd adds the scratch target and the shared sums into v.
int split_finish(const CeedMi355Operator_private *op, void *v, int overwrite, hipStream_t s);
int fill_zero(const CeedMi355Operator_private *op, void *x, hipStream_t s);
// The brick mode's second step: v (=|+=) the partial sums of each brick-surface
// node, in ascending brick order (operator_brick.hip).
int brick_gather(const CeedMi355Operator_private *op, void *v, int overwrite, hipStream_t s);
// v (=|+=) the brick E^T of the element-major E-vector evec[(elem * num_comp +
// comp) * P^3 + node] in packed order: brick sums, then brick_gather.
int brick_evec(const CeedMi355Operator_private *op, const void *evec, void *v, int overwrite, hipStream_t s);

// Kernel launcher for one (P, Q, num_comp, QF, scalar type); nullptr when no instantiation exists.
ApplyFn pick_apply_f64(int P, int Q, int num_comp, int qf);
ApplyFn pick_apply_f32(int P, int Q, int num_comp, int qf);
// Thread layouts of the fused kernel: z-columns with x/y contractions read
// point by point from LDS (hex_apply_kernel), whole 1D lines per thread for
// the x/y contractions (hex_apply_lines_kernel), or a thread per quadrature
// point with all three directions through LDS (hex_apply_points_kernel,
// Q^3 <= 256).
enum { kLayoutColumns = 0, kLayoutLines = 1, kLayoutPoints = 2 };
// Threads an element takes in a layout.
constexpr int layout_threads(int Q, int layout) { return layout == kLayoutPoints ? Q * Q * Q : Q * Q; }
// LDS a gfx950 workgroup may allocate (the CU's whole 160 KiB), and the
// budget the default launch keeps to (so that several workgroups share a CU).
constexpr int kMaxLdsBytes = 160 * 1024, kDefaultLdsBytes = 64 * 1024;
// Threads per workgroup the kernels are compiled for (__launch_bounds__(256);
// 512 was measured: slower line kernels at p = 7, 8 and spills in the large
// column instantiations).
constexpr int max_threads(int) { return 256; }
// Default elements per workgroup: for line-layout diffusion a table of measured
// choices per Q, otherwise as many as fit 256 threads and kDefaultLdsBytes of LDS.
int default_epb(int P, int Q, int qf, int scalar_bytes, int layout);
int lds_bytes(int P, int Q, int qf, int epb, int scalar_bytes, int layout);
// Threads and LDS bytes of the brick kernel's workgroup for a brick of ns
// slots (shape[0] * shape[1] * shape[2]); the kernel's own formulas in
// hex_brick.hpp, for plan set-up on the host.
int64_t brick_threads(int Q, int64_t ns);
int64_t brick_lds_bytes(int P, int Q, int qf, int scalar_bytes, const int *shape);

}  // namespace mi355

namespace mi355 {
// Bricks of face-sharing elements for the brick E^T (operator_brick.hip):
// elements grouped into up to bx * by * bz slots, the packed element order
// (brick by brick, slots ascending) in perm_out, and the device arrays of
// the plan (node lists, partial slots, the gather's map).  Returns
// CEED_MI355_ERROR_UNSUPPORTED (and leaves the plan as it was) where the
// mesh does not form bricks (an element repeats a node).
int build_bricks(CeedMi355Operator_private *op, const int32_t *offsets, int bx, int by, int bz, std::vector<int32_t> &perm_out);
// The host side of build_bricks, without device memory.
struct BrickLayout {
  int                  shape[3] = {1, 1, 1}, nb_nodes = 0;
  int64_t              num_bricks = 0;
  std::vector<int32_t> perm, start, mask, bnode, blist, brow, bslot;
};
int brick_layout(const int32_t *offsets, int64_t num_elem, int P, int num_comp, int bx, int by, int bz, BrickLayout &out);
// Segments of face-sharing elements along x for the segment E^T
// (operator_brick.hip): chains of +x face neighbours cut into balanced pieces
// of at most max_len elements (a piece also ends before an element that would
// share a node with it other than through its -x face), coloured so that no
// two segments of a colour share a node.  Returns CEED_MI355_ERROR_UNSUPPORTED
// where an element repeats a node.
struct SegmentLayout {
  std::vector<int32_t> perm;         // packed position -> caller's element: colour by colour, segment by segment, along x
  std::vector<int32_t> start;        // packed start of each segment in launch order, then the end
  std::vector<int64_t> color;        // colour c owns segments [color[c], color[c+1])
  int                  max_len = 0;  // the longest segment
};
int segment_layout(const int32_t *offsets, int64_t num_elem, int P, int max_len, SegmentLayout &out);
// A plan of CeedMi355OperatorCreateHex with the given kernel launcher and
// geometric-factor components per point (operator_hex.hip).
int create_plan(int qf, int scalar_type, int32_t num_comp, int32_t num_elem, int32_t P_1d, int32_t Q_1d, const double *interp_1d_host,
                const double *grad_1d_host, const int32_t *offsets_host, int32_t comp_stride, int64_t l_size, int flags, int q_comp, ApplyFn fn,
                CeedMi355Operator_private **op_out);
}  // namespace mi355

struct CeedMi355JitKernel;  // run-time compiled fused kernel of a kQfJit plan (operator_jit.hip)
void mi355_jit_kernel_destroy(CeedMi355JitKernel *k);

struct CeedMi355Operator_private {
  int     qf, scalar_type, scalar_bytes, num_comp, P, Q, q_comp;
  int32_t num_elem, comp_stride;
  int64_t l_size;
  int     epb;
  int     layout     = 0;      // mi355::kLayoutColumns or kLayoutLines
  bool    covers_all = false;  // every L entry is reached by some element: Apply needs no zero fill
  int     scatter    = 0;      // CEED_MI355_PLAN_* E^T mode: 0 transpose map, COLOR, SPLIT, ATOMIC, BRICK, SEGMENT
  // transpose offset map (scatter == 0): contribution (elem, comp, node) is
  // stored at d_evec[d_tpos[(elem * num_comp + comp) * P^3 + node]], so that
  // L entry t owns the contiguous run d_evec[d_trow[t] .. d_trow[t+1]), in
  // ascending (elem, comp, node) order
  int32_t *d_trow = nullptr, *d_tpos = nullptr;
  void    *d_evec = nullptr;  // E-vector sorted by L entry, plan scalar type
  // multiplicity split (scatter == CEED_MI355_PLAN_SPLIT): L entries reached by
  // at most two contributions are summed by atomic adds onto a zeroed target
  // (a two-term sum does not depend on the order); the others ("shared",
  // offset bit 0 set) take the transpose map above, restricted to them
  // (d_tpos holds a slot for shared contributions only, d_trow runs for
  // shared entries only); d_srows lists the shared L entries; d_scratch is
  // the zeroed target of ApplyAdd, added to v afterwards
  int32_t *d_srows    = nullptr;
  int64_t  num_shared = 0;
  void    *d_scratch  = nullptr;
  // set by a batched solver: system of each (packed) element, and the
  // solver's per-system active flags; elements of inactive systems are skipped
  const int32_t *d_elem_sys = nullptr, *d_sys_active = nullptr;
  // set by a batched solver with the above: the plan's L range cut into chunks
  // of one system each (begin offsets relative to the plan, the last entry the
  // end); the transpose gather skips the chunks of inactive systems
  const int64_t *d_gather_chunk_begin = nullptr;
  const int32_t *d_gather_chunk_sys   = nullptr;
  int64_t        gather_num_chunks    = 0;
  // brick mode (scatter == CEED_MI355_PLAN_BRICK): brick b owns the packed
  // elements [d_brick_start[b], d_brick_start[b+1]) in the slots set in
  // d_brick_mask[b] (slot s = (s % bx, s / bx % by, s / (bx * by))); its
  // nb_nodes = (bx p + 1)(by p + 1)(bz p + 1) nodes, x fastest, are listed in
  // d_bnode[b] (2 * node + 1 if an element outside the brick shares the node,
  // -1 where no element of the brick lies); partial sums of the shared nodes
  // go to d_partials[b][comp][local]; the gather sums, for each shared node
  // d_blist[j], the partial slots d_bslot[d_brow[j] .. d_brow[j+1]) (each
  // b * num_comp * nb_nodes + local), in ascending brick order
  int      brick_shape[3] = {1, 1, 1};
  int64_t  num_bricks = 0, num_bound = 0;
  int      nb_nodes = 0;
  int32_t *d_brick_start = nullptr, *d_brick_mask = nullptr, *d_bnode = nullptr;
  int32_t *d_blist = nullptr, *d_brow = nullptr, *d_bslot = nullptr;
  void    *d_partials = nullptr;
  // segment mode (scatter == CEED_MI355_PLAN_SEGMENT): segment g owns the
  // packed elements [d_seg_start[g], d_seg_start[g+1]), each sharing its x = 0
  // face with the one before; colour c owns segments [seg_color[c],
  // seg_color[c+1]); one workgroup per segment, seg_max element slots
  int32_t             *d_seg_start = nullptr;
  std::vector<int64_t> seg_color;
  int                  seg_max = 0;
  // colour c owns packed elements [color_start[c], color_start[c+1])
  std::vector<int64_t> color_start;
  std::vector<int32_t> perm;  // packed position -> caller's element index
  int32_t             *d_offsets = nullptr;  // [elem][P^3], packed order
  int32_t             *d_perm    = nullptr;
  void                *d_mats    = nullptr;  // B (Q x P) then D (Q x Q), plan scalar type
  void                *d_qdata   = nullptr;  // [elem][q_comp][Q^3], packed order
  mi355::ApplyFn       apply     = nullptr;
  CeedMi355JitKernel  *jit       = nullptr;  // kQfJit plans: the compiled kernel and the QFunction context
};

namespace mi355 {

template <int P, int Q, int NC, int QF, typename T>
int launch_hex_apply(const CeedMi355Operator_private *op, const void *u, void *v, int overwrite, hipStream_t s) {
  const int    nt     = Q * Q;
  const int    epb    = op->epb;
  const bool   lines  = op->layout == kLayoutLines;
  const bool   points = op->layout == kLayoutPoints;
  const size_t lds    = sizeof(T) * (size_t)(lines ? hex_lines_lds_entries<P, Q, QF, T>(epb)
                                                   : points ? hex_points_lds_entries<P, Q, QF>(epb) : hex_lds_entries<P, Q, QF>(epb));
  if (op->scatter == CEED_MI355_PLAN_BRICK) {
    if (op->num_bricks == 0) return CEED_MI355_SUCCESS;
    const int ns = op->brick_shape[0] * op->brick_shape[1] * op->brick_shape[2];
    hex_apply_brick_kernel<P, Q, NC, QF, T><<<(unsigned)op->num_bricks, BrickThreads<P, Q>::threads(ns), hex_brick_lds_bytes<P, Q, QF, T>(ns, op->nb_nodes), s>>>(
        op->brick_shape[0], op->brick_shape[1], op->brick_shape[2], op->nb_nodes, op->d_brick_start, op->d_brick_mask, op->d_bnode,
        op->comp_stride, (const T *)op->d_qdata, (const T *)op->d_mats, (const T *)u, (T *)v, (T *)op->d_partials, overwrite);
    MI355_LAUNCH_CHECK();
    return brick_gather(op, v, overwrite, s);
  }
  if constexpr (QF == kQfMass || QF == kQfDiffusion) {  // segment plans: mass and diffusion, line layout only
    if (op->scatter == CEED_MI355_PLAN_SEGMENT) {
      for (size_t c = 0; c + 1 < op->seg_color.size(); c++) {
        const int64_t b = op->seg_color[c], e = op->seg_color[c + 1];
        if (e <= b) continue;
        hex_apply_lines_kernel<P, Q, NC, QF, T, true><<<(unsigned)(e - b), epb * nt, lds, s>>>(
            b, e, overwrite ? kScatterOverwrite : kScatterAdd, op->d_offsets, op->comp_stride, (const T *)op->d_qdata, (const T *)op->d_mats,
            (const T *)u, (T *)v, nullptr, nullptr, op->d_elem_sys, op->d_sys_active, op->d_seg_start);
      }
      MI355_LAUNCH_CHECK();
      return CEED_MI355_SUCCESS;
    }
  }
  const bool   tmap  = op->scatter == 0;
  const bool   split = op->scatter == CEED_MI355_PLAN_SPLIT;
  // the split mode's atomic adds land on a zeroed target: v in Apply (zeroed
  // by CeedMi355OperatorApply), the scratch vector in ApplyAdd
  T *target = split && !overwrite ? (T *)op->d_scratch : (T *)v;
  if (split && !overwrite) {
    const int err = fill_zero(op, target, s);
    if (err) return err;
  }
  for (size_t c = 0; c + 1 < op->color_start.size(); c++) {
    const int64_t b = op->color_start[c], e = op->color_start[c + 1];
    if (e <= b) continue;
    const int64_t grid    = (e - b + epb - 1) / epb;
    const int     scatter = tmap ? kScatterEvec
                                 : split ? kScatterSplit
                                         : (op->scatter == CEED_MI355_PLAN_ATOMIC ? kScatterAtomic : (overwrite ? kScatterOverwrite : kScatterAdd));
    if (lines)
      hex_apply_lines_kernel<P, Q, NC, QF, T><<<(unsigned)grid, epb * nt, lds, s>>>(
          b, e, scatter, op->d_offsets, op->comp_stride, (const T *)op->d_qdata, (const T *)op->d_mats, (const T *)u, target, (T *)op->d_evec,
          op->d_tpos, op->d_elem_sys, op->d_sys_active, nullptr);
    else if (points) {
      if constexpr (Q * Q * Q <= 256)
        hex_apply_points_kernel<P, Q, NC, QF, T><<<(unsigned)grid, epb * Q * Q * Q, lds, s>>>(
            b, e, scatter, op->d_offsets, op->comp_stride, (const T *)op->d_qdata, (const T *)op->d_mats, (const T *)u, target,
            (T *)op->d_evec, op->d_tpos, op->d_elem_sys, op->d_sys_active);
      else
        MI355_CHECK(false, CEED_MI355_ERROR_BACKEND, "the point layout needs Q^3 <= 256 (Q = %d)", Q);
    } else
      hex_apply_kernel<P, Q, NC, QF, T><<<(unsigned)grid, epb * nt, lds, s>>>(b, e, scatter, op->d_offsets, op->comp_stride, (const T *)op->d_qdata,
                                                                            (const T *)op->d_mats, (const T *)u, target, (T *)op->d_evec,
                                                                            op->d_tpos, op->d_elem_sys, op->d_sys_active);
  }
  MI355_LAUNCH_CHECK();
  if (tmap) return transpose_gather(op, v, overwrite, s);
  if (split) return split_finish(op, v, overwrite, s);
  return CEED_MI355_SUCCESS;
}

// Dispatch over the instantiated (P, Q) pairs: P = 2..9, Q in {P, P+1}.
template <typename T>
ApplyFn pick_apply(int P, int Q, int nc, int qf) {
#define MI355_CASE_PQ(PP, QQ)                                                                     \
  if (P == PP && Q == QQ) {                                                                       \
    if (qf == kQfMass) return nc == 1 ? launch_hex_apply<PP, QQ, 1, kQfMass, T>                   \
                                      : (nc == 3 ? launch_hex_apply<PP, QQ, 3, kQfMass, T> : nullptr); \
    if (qf == kQfDiffusion) return nc == 1 ? launch_hex_apply<PP, QQ, 1, kQfDiffusion, T>         \
                                           : (nc == 3 ? launch_hex_apply<PP, QQ, 3, kQfDiffusion, T> : nullptr); \
    if (qf == kQfDiffusionKappa) return nc == 1 ? launch_hex_apply<PP, QQ, 1, kQfDiffusionKappa, T> : nullptr; \
    return nullptr;                                                                               \
  }
  MI355_CASE_PQ(2, 2) MI355_CASE_PQ(2, 3)
  MI355_CASE_PQ(3, 3) MI355_CASE_PQ(3, 4)
  MI355_CASE_PQ(4, 4) MI355_CASE_PQ(4, 5)
  MI355_CASE_PQ(5, 5) MI355_CASE_PQ(5, 6)
  MI355_CASE_PQ(6, 6) MI355_CASE_PQ(6, 7)
  MI355_CASE_PQ(7, 7) MI355_CASE_PQ(7, 8)
  MI355_CASE_PQ(8, 8) MI355_CASE_PQ(8, 9)
  MI355_CASE_PQ(9, 9) MI355_CASE_PQ(9, 10)
#undef MI355_CASE_PQ
  return nullptr;
}

}  // namespace mi355
