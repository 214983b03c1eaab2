// Brick and segment E^T for the fused hex plans: grouping the elements into
// bricks of face-sharing elements (host), the gather that finishes the
// brick-surface nodes, the brick E^T of an element-major E-vector (diagonal
// assembly), and the x-segments of the segment E^T (host).
//
// Bricks are found from the element offsets alone (no coordinates): the
// element whose local node (0, 0, 0) is the node at local (P-1, 0, 0) of e,
// and whose whole local x = 0 face equals e's x = P-1 face, is e's +x
// neighbour (likewise +y, +z).  Following the -x links back to an element
// with none gives each element its position along x (and y, z), so that
// bricks start on even positions whatever the element numbering.  Every
// brick is checked node by node: slots whose nodes would contradict the
// brick's node grid stay empty.
#include <string.h>

#include <algorithm>
#include <vector>

#include "hex_plan.hpp"

namespace {

// v[n] (= or +=) sum of the partial slots of brick-surface node n, ascending brick order
template <typename T>
__global__ void __launch_bounds__(256) brick_gather_kernel(int64_t num_bound, int num_comp, int nb_nodes, int32_t comp_stride,
                                                           const int32_t *__restrict__ blist, const int32_t *__restrict__ brow,
                                                           const int32_t *__restrict__ bslot, const T *__restrict__ partials,
                                                           T *__restrict__ v, int overwrite) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < num_bound; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = blist[j];
    const int32_t r0 = brow[j], r1 = brow[j + 1];
    for (int c = 0; c < num_comp; c++) {
      T *dst = v + n + (int64_t)c * comp_stride;
      T  acc = overwrite ? T(0) : *dst;
      for (int32_t m = r0; m < r1; m++) acc += partials[bslot[m] + (int64_t)c * nb_nodes];
      *dst = acc;
    }
  }
}

// Brick E^T of an element-major E-vector evec[(elem * num_comp + comp) * P^3 + node]
// (packed element order): each brick node sums its elements' entries in slot
// order, as the fused brick kernel does, then goes to v or its partial slot.
template <typename T>
__global__ void __launch_bounds__(256) brick_evec_kernel(int P, int bx, int by, int num_comp, int nb_nodes, int32_t comp_stride,
                                                         const int32_t *__restrict__ brick_start, const int32_t *__restrict__ brick_mask,
                                                         const int32_t *__restrict__ bnode, const T *__restrict__ evec,
                                                         T *__restrict__ partials, T *__restrict__ v, int overwrite) {
  const int64_t b  = blockIdx.x;
  const int     p  = P - 1, P3 = P * P * P;
  const int     nbx = bx * p + 1, nby = by * p + 1, bz = (nb_nodes / (nbx * nby) - 1) / p;
  const int64_t e0 = brick_start[b];
  const int     mask = brick_mask[b];
  for (int i = threadIdx.x; i < nb_nodes; i += blockDim.x) {
    const int32_t g2 = bnode[b * nb_nodes + i];
    if (g2 < 0) continue;
    const int bi = i % nbx, bj = (i / nbx) % nby, bk = i / (nbx * nby);
    // slot ranges covering the position along each direction (at most two each)
    const int sx1 = bi / p < bx ? bi / p : bx - 1, sx0 = (bi % p == 0 && bi > 0) ? bi / p - 1 : sx1;
    const int sy1 = bj / p < by ? bj / p : by - 1, sy0 = (bj % p == 0 && bj > 0) ? bj / p - 1 : sy1;
    const int sz1 = bk / p < bz ? bk / p : bz - 1, sz0 = (bk % p == 0 && bk > 0) ? bk / p - 1 : sz1;
    for (int c = 0; c < num_comp; c++) {
      T acc = 0;
      for (int sz = sz0; sz <= sz1; sz++)
        for (int sy = sy0; sy <= sy1; sy++)
          for (int sx = sx0; sx <= sx1; sx++) {
            const int s = sx + bx * (sy + by * sz);
            if (s >= 32 || !((mask >> s) & 1)) continue;
            const int64_t elem = e0 + __builtin_popcount(mask & ((1u << s) - 1));
            const int     x = bi - sx * p, y = bj - sy * p, z = bk - sz * p;
            if (x > p || y > p || z > p) continue;
            acc += evec[(elem * num_comp + c) * P3 + (z * P + y) * P + x];
          }
      if (g2 & 1) partials[(b * num_comp + c) * nb_nodes + i] = acc;
      else {
        T *dst = v + (g2 >> 1) + (int64_t)c * comp_stride;
        *dst   = overwrite ? acc : *dst + acc;
      }
    }
  }
}

// Elements by their local node (0, 0, 0): origin[node] is the element starting
// there, -2 where two elements start at one node, -1 where none does.
std::vector<int32_t> node_origins(const int32_t *offsets, int64_t ne, int P, int64_t n_nodes) {
  const int            es = P * P * P;
  std::vector<int32_t> origin((size_t)n_nodes, -1);
  for (int64_t e = 0; e < ne; e++) {
    int32_t &o = origin[offsets[e * es]];
    o          = o == -1 ? (int32_t)e : -2;
  }
  return origin;
}

// Face neighbours along +d (d = 0, 1, 2 for x, y, z), checked node by node:
// nxt[e] is the element whose whole local face at 0 along d equals e's face at
// P - 1 (-1 if none); prv holds the inverse links (-2 where two elements
// claim one neighbour).
void face_links(const int32_t *offsets, int64_t ne, int P, const std::vector<int32_t> &origin, int d, std::vector<int32_t> &nxt,
                std::vector<int32_t> &prv) {
  const int p = P - 1, es = P * P * P;
  auto      off = [&](int64_t e, int x, int y, int z) { return offsets[e * es + (z * P + y) * P + x]; };
  nxt.assign((size_t)ne, -1);
  prv.assign((size_t)ne, -1);
  for (int64_t e = 0; e < ne; e++) {
    const int32_t c = origin[d == 0 ? off(e, p, 0, 0) : d == 1 ? off(e, 0, p, 0) : off(e, 0, 0, p)];
    if (c < 0 || c == e) continue;
    bool match = true;
    for (int a = 0; a < P && match; a++)
      for (int b2 = 0; b2 < P && match; b2++) {
        if (d == 0) match = off(c, 0, a, b2) == off(e, p, a, b2);
        else if (d == 1) match = off(c, a, 0, b2) == off(e, a, p, b2);
        else match = off(c, a, b2, 0) == off(e, a, b2, p);
      }
    if (!match) continue;
    nxt[e] = c;
    prv[c] = prv[c] == -1 ? (int32_t)e : -2;
  }
}

// Position of every element along the chains of nxt links: steps from the
// start of its chain (an element with no unique predecessor), or from an
// arbitrary element of a chain that closes on itself (periodic meshes).
std::vector<int32_t> chain_positions(const std::vector<int32_t> &nxt, const std::vector<int32_t> &prv) {
  const int64_t        ne = (int64_t)nxt.size();
  std::vector<int32_t> pos((size_t)ne, -1);
  auto                 walk = [&](int64_t e) {
    int32_t k = 0;
    for (int64_t x = e; x >= 0 && pos[x] < 0; x = nxt[x]) pos[x] = k++;
  };
  for (int64_t e = 0; e < ne; e++)
    if (prv[e] < 0) walk(e);
  for (int64_t e = 0; e < ne; e++)
    if (pos[e] < 0) walk(e);
  return pos;
}

}  // namespace

namespace mi355 {

int segment_layout(const int32_t *offsets, int64_t num_elem, int P, int max_len, SegmentLayout &out) {
  const int     es = P * P * P;
  const int64_t ne = num_elem;
  MI355_CHECK(max_len >= 1 && P >= 2, CEED_MI355_ERROR_UNSUPPORTED, "segments of %d elements at P = %d", max_len, P);
  int64_t n_nodes = 0;
  for (int64_t i = 0; i < ne * es; i++) n_nodes = std::max<int64_t>(n_nodes, (int64_t)offsets[i] + 1);
  const std::vector<int32_t> origin = node_origins(offsets, ne, P, n_nodes);
  std::vector<int32_t>       nxt, prv, ynxt, yprv, znxt, zprv;
  face_links(offsets, ne, P, origin, 0, nxt, prv);
  face_links(offsets, ne, P, origin, 1, ynxt, yprv);
  face_links(offsets, ne, P, origin, 2, znxt, zprv);
  const std::vector<int32_t> ypos = chain_positions(ynxt, yprv), zpos = chain_positions(znxt, zprv);
  // chains along +x: from every element without a unique -x neighbour, then
  // from what is left (chains that close on themselves)
  std::vector<uint8_t>              in_chain((size_t)ne, 0);
  std::vector<std::vector<int32_t>> chains;
  auto                              walk = [&](int64_t h) {
    std::vector<int32_t> c;
    for (int64_t x = h, prev = -1; x >= 0 && !in_chain[x] && (prev < 0 || prv[x] == prev); prev = x, x = nxt[x]) {
      c.push_back((int32_t)x);
      in_chain[x] = 1;
    }
    chains.push_back(std::move(c));
  };
  for (int64_t e = 0; e < ne; e++)
    if (prv[e] < 0 && !in_chain[e]) walk(e);
  for (int64_t e = 0; e < ne; e++)
    if (!in_chain[e]) walk(e);
  // chains in (z, y) position order, so that the greedy colouring below meets
  // the segments row by row whatever the element numbering (8 colours on a
  // structured mesh)
  std::stable_sort(chains.begin(), chains.end(), [&](const std::vector<int32_t> &a, const std::vector<int32_t> &b) {
    return std::make_pair(zpos[a[0]], ypos[a[0]]) < std::make_pair(zpos[b[0]], ypos[b[0]]);
  });
  // balanced pieces of at most max_len; a piece also ends before an element
  // that would share a node with it other than through its -x face
  std::vector<std::vector<int32_t>> segs;
  std::vector<int32_t>              stamp((size_t)n_nodes, -1);
  for (const auto &c : chains) {
    const int64_t L = (int64_t)c.size(), k = (L + max_len - 1) / max_len;
    int64_t       at = 0;
    for (int64_t piece = 0; piece < k; piece++) {
      const int64_t len = L / k + (piece < L % k ? 1 : 0), end = at + len;
      while (at < end) {
        const int32_t        id = (int32_t)segs.size();
        std::vector<int32_t> seg;
        for (; at < end; at++) {
          const int32_t e     = c[at];
          const bool    first = seg.empty();
          bool          ok    = true;
          for (int i = 0; i < es && ok; i++) {
            const bool face = !first && i % P == 0;  // the x = 0 face, shared with the element before
            const int32_t g = offsets[(int64_t)e * es + i];
            if (face ? stamp[g] != id : stamp[g] == id) ok = false;
          }
          if (!ok && !first) break;
          MI355_CHECK(ok, CEED_MI355_ERROR_UNSUPPORTED, "element %d repeats a node: no segment E^T for this mesh", e);
          for (int i = 0; i < es; i++) stamp[offsets[(int64_t)e * es + i]] = id;
          seg.push_back(e);
        }
        segs.push_back(std::move(seg));
      }
    }
  }
  // colours: no two segments of a colour share a node (greedy, in segment order)
  std::vector<uint64_t> used_by((size_t)n_nodes, 0);
  std::vector<int32_t>  color(segs.size(), 0);
  int                   num_colors = 0;
  for (size_t g = 0; g < segs.size(); g++) {
    uint64_t used = 0;
    for (int32_t e : segs[g])
      for (int i = 0; i < es; i++) used |= used_by[offsets[(int64_t)e * es + i]];
    MI355_CHECK(~used != 0, CEED_MI355_ERROR_UNSUPPORTED, "segment colouring needs more than 64 colours");
    int k = 0;
    while (used & (1ull << k)) k++;
    color[g]   = k;
    num_colors = std::max(num_colors, k + 1);
    for (int32_t e : segs[g])
      for (int i = 0; i < es; i++) used_by[offsets[(int64_t)e * es + i]] |= 1ull << k;
  }
  // packed order: colour by colour, segments in order, elements along the chain
  out.perm.clear();
  out.perm.reserve((size_t)ne);
  out.start.clear();
  out.color.assign((size_t)num_colors + 1, 0);
  out.max_len = 0;
  for (int k = 0; k < num_colors; k++) {
    for (size_t g = 0; g < segs.size(); g++) {
      if (color[g] != k) continue;
      out.start.push_back((int32_t)out.perm.size());
      out.perm.insert(out.perm.end(), segs[g].begin(), segs[g].end());
      out.max_len = std::max(out.max_len, (int)segs[g].size());
    }
    out.color[k + 1] = (int64_t)out.start.size();
  }
  out.start.push_back((int32_t)out.perm.size());
  MI355_CHECK((int64_t)out.perm.size() == ne, CEED_MI355_ERROR_BACKEND, "segment packing lost elements");
  return CEED_MI355_SUCCESS;
}

int brick_layout(const int32_t *offsets, int64_t num_elem, int P, int num_comp, int bx, int by, int bz, BrickLayout &out) {
  const int     p = P - 1, es = P * P * P, nc = num_comp;
  const int64_t ne = num_elem;
  std::vector<int32_t> &perm_out = out.perm;
  MI355_CHECK(bx >= 1 && by >= 1 && bz >= 1 && bx * by * bz <= 8 && p >= 1, CEED_MI355_ERROR_UNSUPPORTED, "brick shape %dx%dx%d", bx, by, bz);
  auto off = [&](int64_t e, int x, int y, int z) { return offsets[e * es + (z * P + y) * P + x]; };
  int64_t n_nodes = 0;
  for (int64_t i = 0; i < ne * es; i++) n_nodes = std::max<int64_t>(n_nodes, (int64_t)offsets[i] + 1);
  // face neighbours along +x, +y, +z, checked node by node, and the inverse links
  const std::vector<int32_t> origin = node_origins(offsets, ne, P, n_nodes);
  std::vector<int32_t>       nxt[3], prv[3], pos[3];
  for (int d = 0; d < 3; d++) {
    face_links(offsets, ne, P, origin, d, nxt[d], prv[d]);
    pos[d] = chain_positions(nxt[d], prv[d]);
  }
  // bricks
  const int nbx = bx * p + 1, nby = by * p + 1, nbz = bz * p + 1, nb_nodes = nbx * nby * nbz, ns = bx * by * bz;
  std::vector<int32_t> assigned((size_t)ne, 0), stamp((size_t)n_nodes, -1), local((size_t)nb_nodes, -1);
  std::vector<int32_t> brick_slots;  // ns entries per brick, -1 for an empty slot
  std::vector<int32_t> written;
  written.reserve(es);
  auto try_brick = [&](int64_t o) -> bool {
    const int32_t id = (int32_t)(brick_slots.size() / ns);
    std::fill(local.begin(), local.end(), -1);
    int present = 0;
    for (int s = 0; s < ns; s++) {
      const int sx = s % bx, sy = (s / bx) % by, sz = s / (bx * by);
      int64_t   e  = o;
      for (int i = 0; i < sx && e >= 0; i++) e = nxt[0][e];
      for (int i = 0; i < sy && e >= 0; i++) e = nxt[1][e];
      for (int i = 0; i < sz && e >= 0; i++) e = nxt[2][e];
      if (e < 0 || assigned[e]) {
        brick_slots.push_back(-1);
        continue;
      }
      written.clear();
      bool ok = true;
      for (int z = 0; z < P && ok; z++)
        for (int y = 0; y < P && ok; y++)
          for (int x = 0; x < P && ok; x++) {
            const int     li = ((sz * p + z) * nby + (sy * p + y)) * nbx + sx * p + x;
            const int32_t g  = off(e, x, y, z);
            if (local[li] == g) continue;  // shared with an earlier slot, consistently
            if (local[li] != -1 || stamp[g] == id) {
              ok = false;
              break;
            }
            local[li] = g;
            stamp[g]  = id;
            written.push_back(li);
          }
      if (!ok) {
        for (int li : written) {
          stamp[local[li]] = -1;
          local[li]        = -1;
        }
        if (s == 0) {  // the origin itself does not fit: a degenerate element
          brick_slots.resize((size_t)id * ns);
          return false;
        }
        brick_slots.push_back(-1);
        continue;
      }
      assigned[e] = 1;
      present++;
      brick_slots.push_back((int32_t)e);
    }
    return present > 0;
  };
  for (int64_t o = 0; o < ne; o++)
    if (!assigned[o] && pos[0][o] % bx == 0 && pos[1][o] % by == 0 && pos[2][o] % bz == 0) try_brick(o);
  for (int64_t o = 0; o < ne; o++)
    if (!assigned[o] && !try_brick(o)) {
      mi355::set_error("element %lld repeats a node: no brick E^T for this mesh", (long long)o);
      return CEED_MI355_ERROR_UNSUPPORTED;
    }
  const int64_t nbk = (int64_t)(brick_slots.size() / ns);
  // multiplicity of every node (elements containing it)
  std::vector<int32_t> mult((size_t)n_nodes, 0), cnt((size_t)n_nodes, 0);
  for (int64_t i = 0; i < ne * es; i++) mult[offsets[i]]++;
  // packed order, node lists, completeness
  perm_out.clear();
  perm_out.reserve(ne);
  std::vector<int32_t> &start = out.start, &maskv = out.mask, &bnode = out.bnode;
  start.assign((size_t)nbk + 1, 0);
  maskv.assign((size_t)nbk, 0);
  bnode.assign((size_t)(nbk * nb_nodes), -1);
  std::vector<int64_t> bcount((size_t)n_nodes + 1, 0);
  for (int64_t b = 0; b < nbk; b++) {
    start[b] = (int32_t)perm_out.size();
    int m    = 0;
    for (int s = 0; s < ns; s++) {
      const int32_t e = brick_slots[b * ns + s];
      if (e < 0) continue;
      m |= 1 << s;
      perm_out.push_back(e);
      const int sx = s % bx, sy = (s / bx) % by, sz = s / (bx * by);
      for (int z = 0; z < P; z++)
        for (int y = 0; y < P; y++)
          for (int x = 0; x < P; x++) {
            const int li = ((sz * p + z) * nby + (sy * p + y)) * nbx + sx * p + x;
            bnode[b * nb_nodes + li] = off(e, x, y, z);
          }
    }
    maskv[b] = m;
  }
  start[nbk] = (int32_t)perm_out.size();
  // a node is complete in a brick when the brick holds every element containing it
  for (int64_t b = 0; b < nbk; b++) {
    for (int s = 0; s < ns; s++) {
      const int32_t e = brick_slots[b * ns + s];
      if (e < 0) continue;
      for (int i = 0; i < es; i++) cnt[offsets[(int64_t)e * es + i]]++;
    }
    for (int li = 0; li < nb_nodes; li++) {
      int32_t &g2 = bnode[b * nb_nodes + li];
      if (g2 < 0) continue;
      const int32_t g = g2;
      const bool    inc = cnt[g] < mult[g];
      g2              = 2 * g + (inc ? 1 : 0);
      if (inc) bcount[(size_t)g + 1]++;
    }
    for (int s = 0; s < ns; s++) {
      const int32_t e = brick_slots[b * ns + s];
      if (e < 0) continue;
      for (int i = 0; i < es; i++) cnt[offsets[(int64_t)e * es + i]] = 0;
    }
  }
  // the gather's map: brick-surface nodes ascending, their partial slots in brick order
  std::vector<int32_t> &blist = out.blist, &brow = out.brow, &bslot = out.bslot;
  blist.clear();
  brow.assign(1, 0);
  for (int64_t g = 0; g < n_nodes; g++) {
    if (bcount[g + 1] > 0) {
      blist.push_back((int32_t)g);
      brow.push_back(brow.back() + (int32_t)bcount[g + 1]);
    }
    bcount[g + 1] += bcount[g];
  }
  bslot.assign((size_t)bcount[(size_t)n_nodes], 0);
  {
    std::vector<int64_t> cursor(bcount.begin(), bcount.end() - 1);
    for (int64_t b = 0; b < nbk; b++)
      for (int li = 0; li < nb_nodes; li++) {
        const int32_t g2 = bnode[b * nb_nodes + li];
        if (g2 >= 0 && (g2 & 1)) {
          const int64_t slot = b * nc * nb_nodes + li;
          MI355_CHECK(slot < INT32_MAX, CEED_MI355_ERROR_UNSUPPORTED, "brick partials need fewer than 2^31 slots");
          bslot[cursor[g2 >> 1]++] = (int32_t)slot;
        }
      }
  }
  MI355_CHECK((int64_t)perm_out.size() == ne, CEED_MI355_ERROR_BACKEND, "brick packing lost elements");
  out.shape[0]   = bx;
  out.shape[1]   = by;
  out.shape[2]   = bz;
  out.nb_nodes   = nb_nodes;
  out.num_bricks = nbk;
  return CEED_MI355_SUCCESS;
}

int build_bricks(CeedMi355Operator_private *op, const int32_t *offsets, int bx, int by, int bz, std::vector<int32_t> &perm_out) {
  BrickLayout L;
  const int   ierr = brick_layout(offsets, op->num_elem, op->P, op->num_comp, bx, by, bz, L);
  if (ierr) return ierr;
  const int64_t nbk = L.num_bricks;
  op->brick_shape[0] = bx;
  op->brick_shape[1] = by;
  op->brick_shape[2] = bz;
  op->num_bricks     = nbk;
  op->nb_nodes       = L.nb_nodes;
  op->num_bound      = (int64_t)L.blist.size();
  MI355_HIP(hipMalloc(&op->d_brick_start, sizeof(int32_t) * (L.start.size() + 1)));
  MI355_HIP(hipMalloc(&op->d_brick_mask, sizeof(int32_t) * (L.mask.size() + 1)));
  MI355_HIP(hipMalloc(&op->d_bnode, sizeof(int32_t) * (L.bnode.size() + 1)));
  MI355_HIP(hipMalloc(&op->d_blist, sizeof(int32_t) * (L.blist.size() + 1)));
  MI355_HIP(hipMalloc(&op->d_brow, sizeof(int32_t) * L.brow.size()));
  MI355_HIP(hipMalloc(&op->d_bslot, sizeof(int32_t) * (L.bslot.size() + 1)));
  MI355_HIP(hipMalloc(&op->d_partials, (size_t)op->scalar_bytes * ((size_t)nbk * op->num_comp * L.nb_nodes + 1)));
  MI355_HIP(hipMemcpy(op->d_brick_start, L.start.data(), sizeof(int32_t) * L.start.size(), hipMemcpyHostToDevice));
  if (nbk > 0) {
    MI355_HIP(hipMemcpy(op->d_brick_mask, L.mask.data(), sizeof(int32_t) * L.mask.size(), hipMemcpyHostToDevice));
    MI355_HIP(hipMemcpy(op->d_bnode, L.bnode.data(), sizeof(int32_t) * L.bnode.size(), hipMemcpyHostToDevice));
  }
  if (!L.blist.empty()) {
    MI355_HIP(hipMemcpy(op->d_blist, L.blist.data(), sizeof(int32_t) * L.blist.size(), hipMemcpyHostToDevice));
    MI355_HIP(hipMemcpy(op->d_bslot, L.bslot.data(), sizeof(int32_t) * L.bslot.size(), hipMemcpyHostToDevice));
  }
  MI355_HIP(hipMemcpy(op->d_brow, L.brow.data(), sizeof(int32_t) * L.brow.size(), hipMemcpyHostToDevice));
  perm_out = std::move(L.perm);
  return CEED_MI355_SUCCESS;
}

int brick_gather(const CeedMi355Operator_private *op, void *v, int overwrite, hipStream_t s) {
  if (op->num_bound == 0) return CEED_MI355_SUCCESS;
  const int nb = mi355_blocks(op->num_bound, 256, 1 << 20);
  if (op->scalar_type == CEED_MI355_FP64)
    brick_gather_kernel<double><<<nb, 256, 0, s>>>(op->num_bound, op->num_comp, op->nb_nodes, op->comp_stride, op->d_blist, op->d_brow,
                                                   op->d_bslot, (const double *)op->d_partials, (double *)v, overwrite);
  else
    brick_gather_kernel<float><<<nb, 256, 0, s>>>(op->num_bound, op->num_comp, op->nb_nodes, op->comp_stride, op->d_blist, op->d_brow,
                                                  op->d_bslot, (const float *)op->d_partials, (float *)v, overwrite);
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

int brick_evec(const CeedMi355Operator_private *op, const void *evec, void *v, int overwrite, hipStream_t s) {
  if (op->num_bricks > 0) {
    if (op->scalar_type == CEED_MI355_FP64)
      brick_evec_kernel<double><<<(unsigned)op->num_bricks, 256, 0, s>>>(
          op->P, op->brick_shape[0], op->brick_shape[1], op->num_comp, op->nb_nodes, op->comp_stride, op->d_brick_start, op->d_brick_mask,
          op->d_bnode, (const double *)evec, (double *)op->d_partials, (double *)v, overwrite);
    else
      brick_evec_kernel<float><<<(unsigned)op->num_bricks, 256, 0, s>>>(
          op->P, op->brick_shape[0], op->brick_shape[1], op->num_comp, op->nb_nodes, op->comp_stride, op->d_brick_start, op->d_brick_mask,
          op->d_bnode, (const float *)evec, (float *)op->d_partials, (float *)v, overwrite);
    MI355_LAUNCH_CHECK();
  }
  return brick_gather(op, v, overwrite, s);
}

}  // namespace mi355
