// Community locality order: a project-written sequential form of the two
// steps of Rabbit Order (Arai et al., "Rabbit Order: Just-in-Time Parallel
// Reordering for Fast Graph Analysis", IPDPS 2016), the order the reference
// applies before partitioning (rabbit_order_write.py:46-55 running the
// authors' parallel code, rabbit_module/src/rabbit_order.hpp:554-673).
//
//   1. Incremental aggregation.  Vertices are visited once, in increasing
//      degree order (ties by id).  A visited vertex u, the root of the
//      community of the vertices merged into it so far, gathers the edge
//      weight from its community to each neighbouring community c and joins
//      the c with the largest modularity gain
//          dQ(u, c) = 2 (w_uc / 2m - d_u d_c / (2m)^2)
//      (d = total degree of a community, 2m = total degree of the graph) if
//      that gain is positive; otherwise u stays a root.  Edges are gathered
//      lazily: u's own row plus the (community, weight) lists its merged
//      members left with it, each endpoint resolved to its current root.
//   2. Ordering.  A depth-first walk over every root's merge tree, members
//      in merge order, numbers the vertices, so every community and every
//      sub-community occupies a contiguous id range.  Roots are taken in id
//      order, isolated vertices last.
//
// Edge weights are 1 (the structure of the graph, not its values); self
// loops are ignored.  The result depends only on the graph: the visit order,
// the tie rules and the walk are all deterministic.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "cola_internal.hpp"

namespace cola {

std::vector<int32_t> community_order(const Graph &g) {
  const int64_t n = g.n;
  std::vector<int64_t> deg(n, 0);
  int64_t m2 = 0;
  for (int64_t v = 0; v < n; ++v) {
    int64_t d = 0;
    for (int64_t k = g.row_ptr[v]; k < g.row_ptr[v + 1]; ++k) d += g.col_idx[k] != v;
    deg[v] = d;
    m2 += d;
  }
  std::vector<int32_t> visit(n);
  std::iota(visit.begin(), visit.end(), 0);
  std::stable_sort(visit.begin(), visit.end(), [&](int32_t x, int32_t y) { return deg[x] < deg[y]; });

  std::vector<int32_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  std::vector<int64_t> tot(deg);
  std::vector<int32_t> first_child(n, -1), last_child(n, -1), next_sib(n, -1);
  std::vector<uint8_t> visited(n, 0);
  // (community, weight) lists left by merged members with a community that
  // has not been visited yet: chunks of one pool, linked per community.
  struct Chunk {
    int64_t off;
    int64_t len;
    int64_t next;
  };
  std::vector<std::pair<int32_t, int64_t>> pool;
  std::vector<Chunk> chunks;
  std::vector<int64_t> head(n, -1);
  std::vector<int64_t> acc(n, 0);
  std::vector<int32_t> touched;

  auto find = [&](int32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  const double inv = m2 > 0 ? 1.0 / (double)m2 : 0.0;

  for (const int32_t u : visit) {
    visited[u] = 1;
    if (deg[u] == 0) continue;
    touched.clear();
    auto gather = [&](int32_t x, int64_t w) {
      const int32_t r = find(x);
      if (r == u) return;  // inside u's community
      if (acc[r] == 0) touched.push_back(r);
      acc[r] += w;
    };
    for (int64_t k = g.row_ptr[u]; k < g.row_ptr[u + 1]; ++k)
      if (g.col_idx[k] != u) gather(g.col_idx[k], 1);
    for (int64_t c = head[u]; c >= 0; c = chunks[c].next)
      for (int64_t j = chunks[c].off; j < chunks[c].off + chunks[c].len; ++j) gather(pool[j].first, pool[j].second);
    head[u] = -1;

    int32_t best = -1;
    double best_q = 0.0;
    for (const int32_t r : touched) {
      const double q = 2.0 * ((double)acc[r] * inv - (double)tot[u] * (double)tot[r] * inv * inv);
      if (q > best_q) {  // first maximum in gather order
        best_q = q;
        best = r;
      }
    }
    if (best >= 0) {
      parent[u] = best;
      tot[best] += tot[u];
      if (last_child[best] < 0) first_child[best] = u;
      else next_sib[last_child[best]] = u;
      last_child[best] = u;
      if (!visited[best]) {
        // The rest of u's edges become edges of best's community, needed
        // when best is visited.
        Chunk ch{(int64_t)pool.size(), 0, head[best]};
        for (const int32_t r : touched)
          if (r != best) {
            pool.emplace_back(r, acc[r]);
            ++ch.len;
          }
        if (ch.len > 0) {
          chunks.push_back(ch);
          head[best] = (int64_t)chunks.size() - 1;
        }
      }
    }
    for (const int32_t r : touched) acc[r] = 0;
  }
  std::vector<std::pair<int32_t, int64_t>>().swap(pool);
  std::vector<Chunk>().swap(chunks);

  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> stack, kids;
  for (int64_t v = 0; v < n; ++v) {
    if (parent[v] != v || (deg[v] == 0 && first_child[v] < 0)) continue;
    stack.push_back((int32_t)v);
    while (!stack.empty()) {
      const int32_t x = stack.back();
      stack.pop_back();
      order.push_back(x);
      kids.clear();
      for (int32_t c = first_child[x]; c >= 0; c = next_sib[c]) kids.push_back(c);
      stack.insert(stack.end(), kids.rbegin(), kids.rend());  // popped in merge order
    }
  }
  for (int64_t v = 0; v < n; ++v)
    if (deg[v] == 0 && first_child[v] < 0) order.push_back((int32_t)v);
  return order;
}

}  // namespace cola
