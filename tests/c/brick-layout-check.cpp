// CPU check of the brick E^T layout (mi355::brick_layout, the host half of
// the brick plan): builds a structured hex mesh (optionally with shuffled
// elements, renumbered nodes with unused L entries, odd element counts), runs
// the layout, and checks
//   * the packed order is a permutation of the elements;
//   * every brick's node list agrees with the element offsets at every slot;
//   * a node is marked complete exactly when the brick holds every element
//     containing it;
//   * the gather map lists every incomplete (brick, position) once, grouped by
//     node, in ascending brick order;
//   * E^T of random element values through the layout (brick sums in slot
//     order, then partial sums in brick order) equals a direct sum to 1e-13.
// Usage: brick-layout-check n p bx by bz shuffle renumber
// Prints "ok bricks=... fill=... bound=..." and exits 0, or a message and 1.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "hex_plan.hpp"

static int fail(const char *msg, long long a = 0, long long b = 0) {
  printf("FAIL: %s (%lld, %lld)\n", msg, a, b);
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 8) return fail("usage");
  const int n = atoi(argv[1]), p = atoi(argv[2]), bx = atoi(argv[3]), by = atoi(argv[4]), bz = atoi(argv[5]);
  const int shuffle = atoi(argv[6]), renumber = atoi(argv[7]);
  const int P = p + 1, es = P * P * P, nn = n * p + 1;
  const int64_t ne = (int64_t)n * n * n;
  std::mt19937_64 rng(1234);
  // node numbering: lexicographic, or a random injective map into a larger range
  std::vector<int32_t> nodemap((size_t)nn * nn * nn);
  std::iota(nodemap.begin(), nodemap.end(), 0);
  int64_t n_nodes = (int64_t)nodemap.size();
  if (renumber) {
    std::vector<int32_t> all((size_t)n_nodes + n_nodes / 3);
    std::iota(all.begin(), all.end(), 0);
    std::shuffle(all.begin(), all.end(), rng);
    for (size_t i = 0; i < nodemap.size(); i++) nodemap[i] = all[i];
    n_nodes = (int64_t)all.size();
  }
  std::vector<int64_t> eorder((size_t)ne);
  std::iota(eorder.begin(), eorder.end(), 0);
  if (shuffle) std::shuffle(eorder.begin(), eorder.end(), rng);
  std::vector<int32_t> off((size_t)(ne * es));
  for (int64_t k = 0; k < ne; k++) {
    const int64_t e = eorder[k];
    const int     ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((int64_t)n * n));
    for (int z = 0; z < P; z++)
      for (int y = 0; y < P; y++)
        for (int x = 0; x < P; x++)
          off[k * es + (z * P + y) * P + x] = nodemap[((int64_t)(ez * p + z) * nn + (ey * p + y)) * nn + (ex * p + x)];
  }
  mi355::BrickLayout L;
  const int nc = 1;
  if (mi355::brick_layout(off.data(), ne, P, nc, bx, by, bz, L)) return fail("brick_layout returned an error");
  const int     NB = L.nb_nodes, ns = bx * by * bz, nbx = bx * p + 1, nby = by * p + 1;
  const int64_t nbk = L.num_bricks;
  if (NB != nbx * nby * (bz * p + 1)) return fail("nb_nodes", NB);
  // permutation
  if ((int64_t)L.perm.size() != ne) return fail("perm size", (long long)L.perm.size(), ne);
  {
    std::vector<char> seen((size_t)ne, 0);
    for (int32_t e : L.perm) {
      if (e < 0 || e >= ne || seen[e]) return fail("perm not a permutation", e);
      seen[e] = 1;
    }
  }
  if ((int64_t)L.start.size() != nbk + 1 || L.start[0] != 0 || L.start[nbk] != ne) return fail("brick starts");
  std::vector<int32_t> mult((size_t)n_nodes, 0);
  for (int32_t o : off) mult[o]++;
  // node lists against the offsets; completeness
  int64_t present = 0;
  for (int64_t b = 0; b < nbk; b++) {
    const int mask = L.mask[b];
    if (mask <= 0 || mask >= (1 << ns)) return fail("mask", b, mask);
    if (__builtin_popcount(mask) != L.start[b + 1] - L.start[b]) return fail("mask/start mismatch", b);
    present += __builtin_popcount(mask);
    std::vector<int> cover((size_t)NB, 0);
    int              k = 0;
    for (int s = 0; s < ns; s++) {
      if (!((mask >> s) & 1)) continue;
      const int32_t e  = L.perm[L.start[b] + k++];
      const int     sx = s % bx, sy = (s / bx) % by, sz = s / (bx * by);
      for (int z = 0; z < P; z++)
        for (int y = 0; y < P; y++)
          for (int x = 0; x < P; x++) {
            const int     li = ((sz * p + z) * nby + (sy * p + y)) * nbx + sx * p + x;
            const int32_t g2 = L.bnode[b * NB + li];
            if (g2 < 0 || (g2 >> 1) != off[(int64_t)e * es + (z * P + y) * P + x]) return fail("node list differs from offsets", b, li);
            cover[li]++;
          }
    }
    for (int li = 0; li < NB; li++) {
      const int32_t g2 = L.bnode[b * NB + li];
      if (!cover[li]) {
        if (g2 != -1) return fail("uncovered position holds a node", b, li);
        continue;
      }
      const bool complete = cover[li] == mult[g2 >> 1];
      if (complete == (bool)(g2 & 1)) return fail("completeness bit", b, li);
    }
  }
  // gather map
  std::vector<int64_t> seen_slot;
  {
    int64_t prev_node = -1;
    if ((int64_t)L.brow.size() != (int64_t)L.blist.size() + 1 || L.brow[0] != 0) return fail("brow size");
    for (size_t j = 0; j < L.blist.size(); j++) {
      if (L.blist[j] <= prev_node) return fail("blist not ascending", j);
      prev_node = L.blist[j];
      int64_t prev_b = -1;
      for (int32_t m = L.brow[j]; m < L.brow[j + 1]; m++) {
        const int64_t slot = L.bslot[m], b = slot / (nc * NB), li = slot % (nc * NB);
        if (b <= prev_b || b >= nbk) return fail("bslot not in ascending brick order", j, m);
        prev_b           = b;
        const int32_t g2 = L.bnode[b * NB + li];
        if (g2 < 0 || !(g2 & 1) || (g2 >> 1) != L.blist[j]) return fail("bslot points at the wrong node", j, m);
        seen_slot.push_back(slot);
      }
    }
    int64_t incomplete = 0;
    for (int32_t g2 : L.bnode) incomplete += g2 >= 0 && (g2 & 1);
    std::sort(seen_slot.begin(), seen_slot.end());
    if ((int64_t)seen_slot.size() != incomplete || std::adjacent_find(seen_slot.begin(), seen_slot.end()) != seen_slot.end())
      return fail("gather map does not list every incomplete position once", (long long)seen_slot.size(), incomplete);
  }
  // E^T of random element values: direct sum against the layout's two-level sum
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<double>                    ev((size_t)(ne * es));
  for (double &x : ev) x = U(rng);
  std::vector<double> ref((size_t)n_nodes, 0.0), got((size_t)n_nodes, 0.0), part((size_t)(nbk * NB), 0.0);
  for (int64_t i = 0; i < ne * es; i++) ref[off[i]] += ev[i];
  std::vector<int64_t> pos_of((size_t)ne);  // caller element -> packed position
  for (int64_t k = 0; k < ne; k++) pos_of[L.perm[k]] = k;
  for (int64_t b = 0; b < nbk; b++) {
    std::vector<double> acc((size_t)NB, 0.0);
    int                 k = 0;
    for (int s = 0; s < ns; s++) {
      if (!((L.mask[b] >> s) & 1)) continue;
      const int32_t e  = L.perm[L.start[b] + k++];
      const int     sx = s % bx, sy = (s / bx) % by, sz = s / (bx * by);
      for (int z = 0; z < P; z++)
        for (int y = 0; y < P; y++)
          for (int x = 0; x < P; x++)
            acc[((sz * p + z) * nby + (sy * p + y)) * nbx + sx * p + x] += ev[(int64_t)e * es + (z * P + y) * P + x];
    }
    for (int li = 0; li < NB; li++) {
      const int32_t g2 = L.bnode[b * NB + li];
      if (g2 < 0) continue;
      if (g2 & 1) part[b * NB + li] = acc[li];
      else got[g2 >> 1] = acc[li];
    }
  }
  for (size_t j = 0; j < L.blist.size(); j++) {
    double a = 0;
    for (int32_t m = L.brow[j]; m < L.brow[j + 1]; m++) a += part[L.bslot[m]];
    got[L.blist[j]] = a;
  }
  double num = 0, den = 0;
  for (int64_t i = 0; i < n_nodes; i++) num += (got[i] - ref[i]) * (got[i] - ref[i]), den += ref[i] * ref[i];
  if (!(sqrt(num) <= 1e-13 * sqrt(den))) {
    printf("FAIL: E^T mismatch rel %.3e\n", sqrt(num / den));
    return 1;
  }
  printf("ok bricks=%lld fill=%.4f bound=%lld partials=%lld\n", (long long)nbk, (double)present / (double)(nbk * ns), (long long)L.blist.size(),
         (long long)seen_slot.size());
  return 0;
}
