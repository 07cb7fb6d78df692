// Graph-walk corpus: DeepWalk-style sentences (uniform random walks) over a
// Barabasi-Albert preferential-attachment graph.
//
//  * ba_graph (host): nodes 0..m form a complete graph; every later node v draws m
//    distinct targets, each uniform over the list of edge endpoints as it stands before
//    v's edges (so a node is drawn in proportion to its degree), draw i from the keyed
//    stream key(seed, (0x42414752 << 32) ^ v, i). Edges are kept in creation order and
//    the adjacency lists (CSR, both directions) in edge order. The construction is
//    sequential by nature (each node's draws depend on every earlier edge) and takes
//    well under a second for a million nodes.
//  * walks (device): walk w = r*n + u starts at node u; x0 = key(seed, (0x57414C4B << 32) ^ w,
//    0) truncates it, when (x0 >> 32) < thr, to a length uniform in [lmin, L]; step s
//    moves to neighbor ((key(.., w, s) >> 32) * deg) >> 32 of the current node; node u is
//    written as token base + u. One thread per walk: a lengths kernel, an exclusive scan
//    into sentence offsets, a tokens kernel.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "common.h"

namespace w2v {

constexpr uint64_t kSaltBA = 0x42414752ull << 32;
constexpr uint64_t kSaltWalk = 0x57414C4Bull << 32;

int64_t ba_edge_count(int64_t n, int32_t m) {
  if (m < 1 || n <= m) return -1;
  return (int64_t)m * (m + 1) / 2 + (n - m - 1) * (int64_t)m;
}

int64_t ba_graph(int64_t n, int32_t m, uint64_t seed, int64_t* off, int32_t* adj) {
  const int64_t E = ba_edge_count(n, m);
  if (E < 0 || n > INT32_MAX) return -1;
  std::vector<int32_t> ea((size_t)E), eb((size_t)E), rep;
  rep.reserve((size_t)E * 2);
  int64_t e = 0;
  for (int32_t i = 0; i <= m; i++)
    for (int32_t j = i + 1; j <= m; j++) {
      ea[e] = i;
      eb[e] = j;
      e++;
      rep.push_back(i);
      rep.push_back(j);
    }
  const uint64_t ms = mix64(seed);
  std::vector<int32_t> chosen((size_t)m);
  for (int64_t v = m + 1; v < n; v++) {
    const uint64_t len = rep.size();
    int32_t got = 0;
    uint64_t attempt = 0;
    while (got < m) {
      const uint64_t x = key2(ms, kSaltBA ^ (uint64_t)v, attempt++);
      const int32_t t = rep[((x >> 32) * len) >> 32];
      bool dup = false;
      for (int32_t k = 0; k < got; k++) dup = dup || chosen[k] == t;
      if (!dup) chosen[got++] = t;
    }
    for (int32_t k = 0; k < m; k++) {
      ea[e] = (int32_t)v;
      eb[e] = chosen[k];
      e++;
      rep.push_back((int32_t)v);
      rep.push_back(chosen[k]);
    }
  }
  for (int64_t i = 0; i <= n; i++) off[i] = 0;
  for (int64_t k = 0; k < E; k++) {
    off[ea[k] + 1]++;
    off[eb[k] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) off[i + 1] += off[i];
  std::vector<int64_t> cur(off, off + n);
  for (int64_t k = 0; k < E; k++) {
    adj[cur[ea[k]]++] = eb[k];
    adj[cur[eb[k]]++] = ea[k];
  }
  return E;
}

__device__ __forceinline__ int64_t walk_length(uint64_t ms, int64_t w, int32_t L, int32_t lmin, uint64_t thr) {
  const uint64_t x0 = key2(ms, kSaltWalk ^ (uint64_t)w, 0);
  if ((x0 >> 32) < thr) return lmin + (int64_t)(((x0 & 0xFFFFFFFFull) * (uint64_t)(L - lmin + 1)) >> 32);
  return L;
}

// len[w] for w < n_walks and len[n_walks] = 0 (the scan's last entry)
__global__ void walk_len_kernel(int64_t n_walks, uint64_t ms, int32_t L, int32_t lmin, uint64_t thr,
                                int64_t* __restrict__ len) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < n_walks) len[w] = walk_length(ms, w, L, lmin, thr);
  else if (w == n_walks) len[w] = 0;
}

__global__ void walk_tokens_kernel(int64_t n_walks, int64_t n_nodes, uint64_t ms,
                                   const int64_t* __restrict__ goff, const int32_t* __restrict__ adj,
                                   const int64_t* __restrict__ soff, int32_t base,
                                   int32_t* __restrict__ out) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_walks) return;
  int32_t cur = (int32_t)(w % n_nodes);
  const int64_t b = soff[w], len = soff[w + 1] - b;
  out[b] = base + cur;
  for (int64_t s = 1; s < len; s++) {
    const int64_t o = goff[cur], deg = goff[cur + 1] - o;
    if (deg > 0) {
      const uint64_t x = key2(ms, kSaltWalk ^ (uint64_t)w, (uint64_t)s);
      cur = adj[o + (int64_t)(((x >> 32) * (uint64_t)deg) >> 32)];
    }
    out[b + s] = base + cur;
  }
}

hipError_t launch_walk_lengths(int64_t n_walks, uint64_t ms, int32_t L, int32_t lmin, uint64_t thr,
                               int64_t* len, hipStream_t st) {
  hipLaunchKernelGGL(walk_len_kernel, dim3((unsigned)((n_walks + 1 + 255) / 256)), dim3(256), 0, st,
                     n_walks, ms, L, lmin, thr, len);
  return hipGetLastError();
}

hipError_t launch_walk_tokens(int64_t n_walks, int64_t n_nodes, uint64_t ms, const int64_t* goff,
                              const int32_t* adj, const int64_t* soff, int32_t base, int32_t* out,
                              hipStream_t st) {
  hipLaunchKernelGGL(walk_tokens_kernel, dim3((unsigned)((n_walks + 255) / 256)), dim3(256), 0, st,
                     n_walks, n_nodes, ms, goff, adj, soff, base, out);
  return hipGetLastError();
}

}  // namespace w2v
