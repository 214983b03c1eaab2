// CPU check of the segment E^T layout (mi355::segment_layout, the host half of
// the segment plan): builds a structured hex mesh (optionally with shuffled
// elements and nodes renumbered into an L-vector with unused entries), runs the
// layout, and checks
//   * the packed order is a permutation of the elements, segments cover it in
//     order, each of 1..max_len elements, colours cover the segments in order;
//   * consecutive elements of a segment share exactly one face: the x = 0 face
//     of the later one is the x = P-1 face of the earlier one, node for node,
//     and no other node is shared inside a segment;
//   * no two segments of one colour share a node;
//   * E^T of random element values the way the segment kernel forms it (each
//     face between consecutive elements summed left value first and written by
//     the left element, colour after colour, the first write of a node
//     overwriting) equals a direct sum to 1e-13;
//   * on a structured mesh every x-row of n elements is cut into
//     ceil(n / max_len) segments of balanced lengths.
// Usage: segment-layout-check n p max_len shuffle renumber
// Prints "ok segments=... colors=... mean_len=..." and exits 0, or a message and 1.
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
  if (argc < 6) return fail("usage");
  const int n = atoi(argv[1]), p = atoi(argv[2]), max_len = atoi(argv[3]), shuffle = atoi(argv[4]), renumber = atoi(argv[5]);
  const int     P = p + 1, es = P * P * P, nn = n * p + 1;
  const int64_t ne = (int64_t)n * n * n;
  std::mt19937_64 rng(4321);
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
  auto at = [&](int64_t e, int x, int y, int z) { return off[e * es + (z * P + y) * P + x]; };

  mi355::SegmentLayout L;
  if (mi355::segment_layout(off.data(), ne, P, max_len, L)) return fail("segment_layout returned an error");
  // packed order, segments, colours
  if ((int64_t)L.perm.size() != ne) return fail("perm size", (long long)L.perm.size(), ne);
  {
    std::vector<int32_t> s = L.perm;
    std::sort(s.begin(), s.end());
    for (int64_t i = 0; i < ne; i++)
      if (s[i] != i) return fail("perm is not a permutation", i, s[i]);
  }
  const int64_t ns = (int64_t)L.start.size() - 1;
  if (ns < 0 || L.start[0] != 0 || L.start[ns] != ne) return fail("segment starts do not cover the elements");
  int longest = 0;
  for (int64_t g = 0; g < ns; g++) {
    const int len = L.start[g + 1] - L.start[g];
    if (len < 1 || len > max_len) return fail("segment length", g, len);
    longest = std::max(longest, len);
  }
  if (longest != L.max_len) return fail("max_len", longest, L.max_len);
  const int nc = (int)L.color.size() - 1;
  if (nc < 0 || L.color[0] != 0 || L.color[nc] != ns) return fail("colours do not cover the segments");
  for (int c = 0; c < nc; c++)
    if (L.color[c + 1] < L.color[c]) return fail("colour ranges", c);
  // faces inside segments; nothing else shared inside a segment
  std::vector<int64_t> stamp((size_t)n_nodes, -1);
  for (int64_t g = 0; g < ns; g++)
    for (int32_t k = L.start[g]; k < L.start[g + 1]; k++) {
      const int64_t e = L.perm[k];
      for (int z = 0; z < P; z++)
        for (int y = 0; y < P; y++)
          for (int x = 0; x < P; x++) {
            const int32_t node = at(e, x, y, z);
            if (k > L.start[g] && x == 0) {
              if (node != at(L.perm[k - 1], p, y, z)) return fail("x-face does not match the element before", g, k);
              continue;
            }
            if (stamp[node] == g) return fail("node shared inside a segment off the chained faces", g, node);
            stamp[node] = g;
          }
    }
  // colours: no shared node between two segments of a colour
  std::fill(stamp.begin(), stamp.end(), -1);
  for (int c = 0; c < nc; c++)
    for (int64_t g = L.color[c]; g < L.color[c + 1]; g++) {
      std::vector<int32_t> nodes;
      for (int32_t k = L.start[g]; k < L.start[g + 1]; k++)
        for (int i = 0; i < es; i++) nodes.push_back(off[(int64_t)L.perm[k] * es + i]);
      std::sort(nodes.begin(), nodes.end());
      nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
      for (int32_t node : nodes) {
        if (stamp[node] >= L.color[c] && stamp[node] != g) return fail("two segments of one colour share a node", c, node);
        stamp[node] = g;
      }
    }
  // E^T of random element values the segment kernel's way against a direct sum
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<double>                    ev((size_t)(ne * es)), direct((size_t)n_nodes, 0.0), v((size_t)n_nodes, 7.0);
  for (auto &x : ev) x = U(rng);
  for (int64_t e = 0; e < ne; e++)
    for (int i = 0; i < es; i++) direct[off[e * es + i]] += ev[e * es + i];
  std::vector<char> touched((size_t)n_nodes, 0);
  for (int64_t g = 0; g < ns; g++)
    for (int32_t k = L.start[g]; k < L.start[g + 1]; k++) {
      const int64_t e = L.perm[k];
      for (int z = 0; z < P; z++)
        for (int y = 0; y < P; y++)
          for (int x = 0; x < P; x++) {
            if (k > L.start[g] && x == 0) continue;  // written by the element before
            double s = ev[e * es + (z * P + y) * P + x];
            if (x == p && k + 1 < L.start[g + 1]) s += ev[(int64_t)L.perm[k + 1] * es + (z * P + y) * P];
            const int32_t node = at(e, x, y, z);
            v[node]            = touched[node] ? v[node] + s : s;
            touched[node]      = 1;
          }
    }
  double err = 0, ref = 0;
  for (int64_t i = 0; i < n_nodes; i++) {
    if (!touched[i]) continue;
    err = std::max(err, fabs(v[i] - direct[i]));
    ref = std::max(ref, fabs(direct[i]));
  }
  if (err > 1e-13 * std::max(1.0, ref)) return fail("E^T through the segments differs from the direct sum");
  // structured rows: ceil(n / max_len) balanced segments each
  const int64_t k_row = (n + max_len - 1) / max_len;
  if (ns != k_row * n * n) return fail("segments per x-row", ns, k_row * n * n);
  for (int64_t g = 0; g < ns; g++) {
    const int len = L.start[g + 1] - L.start[g];
    if (len != n / k_row && len != n / k_row + 1) return fail("unbalanced segment", g, len);
  }
  printf("ok segments=%lld colors=%d mean_len=%.3f\n", (long long)ns, nc, ns ? (double)ne / ns : 0.0);
  return 0;
}
