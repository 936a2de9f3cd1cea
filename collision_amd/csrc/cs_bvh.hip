// Broad phase: element boxes and a linear BVH (LBVH) for gfx950.
//
// Replaces the reference's incremental AABB tree (AABB.cpp:129-343).  The
// reference's candidate set is "every unordered pair of distinct elements
// whose boxes, inflated by 1e-6, overlap under the closed test"
// (AABB.cpp:10-33, 56-60, 284-343); the tree only fixes the visiting order.
// Here:
//   * k_element_boxes : FP64 static or swept boxes per element (the
//     reference's min/max_{static,moving}_coord, dcollid.cpp:852-932).
//   * k_bounds / k_morton : 30-bit Morton codes of FP32 box centres.
//   * (radix sort of codes, cs_sort.hip)
//   * k_karras        : radix tree over the sorted codes (index tie-break);
//     k_preorder / k_permute : the tree laid out in depth-first pre-order
//     with escape links (ropes) for stackless traversal.
//   * k_refit_tiles / k_refit_top : bottom-up FP32 boxes rounded outward,
//     so every FP64 overlap is also an FP32 overlap.
//   * k_traverse      : one lane per leaf in Morton order; it starts at its
//     own leaf's rope, so it visits exactly the subtrees after itself in
//     depth-first order (every pair once, no stack); a
//     leaf overlap is confirmed with the exact FP64 boxes, then the
//     reference's early rejects (shared vertex, same rigid surface;
//     dcollid.cpp:762, 805; dcollid3d.cpp:209-214, 257-264, 279-284) run
//     before the pair is appended to the compacted candidate stream.
#include <algorithm>

#include "cs_internal.h"

namespace cs {

__global__ void k_element_boxes(const int4* __restrict__ ev, const dvec4* __restrict__ x,
                                const dvec4* __restrict__ x_old, const dvec4* __restrict__ av,
                                double dt, int moving, int ne, dbox* __restrict__ ebox) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int4 t = ev[e];
  const int np = (t.z >= 0) ? 3 : 2;
  const int vid[3] = {t.x, t.y, t.z};
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
  double hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = 0; i < np; ++i) {
    const int v = vid[i];
    if (moving) {
      const dvec4 xo = x_old[v];
      const dvec4 a = av[v];
      const double o[3] = {xo.x, xo.y, xo.z};
      const double n[3] = {xo.x + a.x * dt, xo.y + a.y * dt, xo.z + a.z * dt};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        hi[d] = smax(hi[d], o[d]);
        hi[d] = smax(hi[d], n[d]);
        lo[d] = smin(lo[d], o[d]);
        lo[d] = smin(lo[d], n[d]);
      }
    } else {
      // triangles use x_old, bonds use Coords (equal to x_old in the
      // proximity pass; dcollid.cpp:852-860, 892-910)
      const dvec4 p = (np == 3) ? x_old[v] : x[v];
      const double o[3] = {p.x, p.y, p.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        hi[d] = smax(o[d], hi[d]);
        lo[d] = smin(o[d], lo[d]);
      }
    }
  }
  dbox b;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = lo[d] - CS_BOX_EPS;
    b.hi[d] = hi[d] + CS_BOX_EPS;
  }
  ebox[e] = b;
}

// min/max of box centres (FP32) -> bounds[0..5]; one atomic per block
__global__ void k_bounds(const dbox* __restrict__ ebox, int ne, float* __restrict__ bounds) {
  __shared__ float s[6][256];
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    const dbox b = ebox[e];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float c = (float)(0.5 * (b.lo[d] + b.hi[d]));
      mn[d] = fminf(mn[d], c);
      mx[d] = fmaxf(mx[d], c);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s[d][threadIdx.x] = mn[d];
    s[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + off]);
        s[3 + d][threadIdx.x] = fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + off]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      atomicMin_float(&bounds[d], s[d][0]);
      atomicMax_float(&bounds[3 + d], s[3 + d][0]);
    }
  }
}

__device__ inline uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_morton(const dbox* __restrict__ ebox, int ne, const float* __restrict__ bounds,
                         uint32_t* __restrict__ codes, uint32_t* __restrict__ ids) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const dbox b = ebox[e];
  uint32_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float c = (float)(0.5 * (b.lo[d] + b.hi[d]));
    float ext = bounds[3 + d] - bounds[d];
    float u = ext > 0.f ? (c - bounds[d]) / ext : 0.5f;
    u = fminf(fmaxf(u, 0.f), 1.f);
    q[d] = (uint32_t)fminf(u * 1024.f, 1023.f);
  }
  codes[e] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  ids[e] = (uint32_t)e;
}

__device__ inline int lcp_delta(const uint32_t* __restrict__ codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = codes[i], b = codes[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

// Karras (2012) radix tree, built in construction arrays: internal nodes
// 0..n-2, leaf j is node n-1+j.  The node splitting after leaf g records its
// second child in split_right[g]: that child is what depth-first order
// visits right after any subtree whose last leaf is g (every split position
// 0..n-2 belongs to exactly one internal node).
__global__ void k_karras(const uint32_t* __restrict__ codes, int n, int32_t* __restrict__ kleft,
                         int32_t* __restrict__ kright, int32_t* __restrict__ kparent,
                         int32_t* __restrict__ split_right, int32_t* __restrict__ krmin,
                         int32_t* __restrict__ krmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (lcp_delta(codes, n, i, i + 1) - lcp_delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = lcp_delta(codes, n, i, i - d);
  int lmax = 2;
  while (lcp_delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (lcp_delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = lcp_delta(codes, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (lcp_delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const int left = (lo == gamma) ? (n - 1) + gamma : gamma;
  const int right = (hi == gamma + 1) ? (n - 1) + gamma + 1 : gamma + 1;
  kleft[i] = left;
  kright[i] = right;
  split_right[gamma] = right;
  krmin[i] = lo;
  krmax[i] = hi;
  kparent[left] = i;
  kparent[right] = i;
  if (i == 0) kparent[0] = -1;
}

// Depth-first (pre-order) position of every construction node: the nodes
// before x are its ancestors and the subtrees left of its path, which hold
// exactly the leaves before x's first leaf a, so pre(x) = 2a + (number of
// left turns from the root to x).  The traversal layout is pre-order: a
// node's first child is the next record and its rope points forward, so a
// walk through the tree touches memory in increasing address order.
__global__ void k_preorder(int n, const int32_t* __restrict__ kleft,
                           const int32_t* __restrict__ kparent, const int32_t* __restrict__ krmin,
                           int32_t* __restrict__ pre) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * n - 1) return;
  const int a = (x < n - 1) ? krmin[x] : x - (n - 1);
  int left_turns = 0;
  for (int c = x, p = kparent[x]; p >= 0; c = p, p = kparent[p])
    if (kleft[p] == c) ++left_turns;
  pre[x] = 2 * a + left_turns;
}

// scatter the construction tree into the pre-order arrays
__global__ void k_permute(int n, const int32_t* __restrict__ pre, const int32_t* __restrict__ kright,
                          const int32_t* __restrict__ kparent, const int32_t* __restrict__ split_right,
                          const int32_t* __restrict__ krmin, const int32_t* __restrict__ krmax,
                          const uint32_t* __restrict__ sorted_ids, bvh_node* __restrict__ nodes,
                          int32_t* __restrict__ parent, int32_t* __restrict__ nright,
                          int32_t* __restrict__ rmin, int32_t* __restrict__ rmax,
                          int32_t* __restrict__ lpre, int32_t* __restrict__ enode) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * n - 1) return;
  const int p = pre[x];
  const bool internal = x < n - 1;
  const int last = internal ? krmax[x] : x - (n - 1);
  const int rk = (last == n - 1) ? -1 : split_right[last];
  nodes[p].rope = rk < 0 ? -1 : pre[rk];
  parent[p] = kparent[x] < 0 ? -1 : pre[kparent[x]];
  if (internal) {
    nodes[p].left = -1;
    nright[p] = pre[kright[x]];
    rmin[p] = krmin[x];
    rmax[p] = krmax[x];
  } else {
    const int j = x - (n - 1);
    const int e = (int)sorted_ids[j];
    nodes[p].left = e;
    nright[p] = -1;
    rmin[p] = j;
    rmax[p] = j;
    lpre[j] = p;
    enode[e] = p;
  }
}

// Refit.  Every internal node covers a contiguous range [rmin, rmax] of
// Morton-sorted leaves.  Nodes whose range lies inside one tile of
// REFIT_TILE leaves are refit by that tile's work-group, bottom-up: the
// second lane to arrive at a node (work-group-scope atomic counter) merges
// both children.  The remaining "top" nodes are refit the same way by a
// single work-group, starting from the top nodes whose children are all
// tile-local.  No agent-scope fence is needed: every counter is touched by
// one work-group only, and data crosses work-groups only at kernel
// boundaries.  Leaf boxes are the FP64 boxes rounded outward to FP32.
constexpr int REFIT_THREADS = 1024;
constexpr int REFIT_TILE = 8192;


__device__ inline bool is_top(const int32_t* rmin, const int32_t* rmax, int node) {
  return (rmin[node] / REFIT_TILE) != (rmax[node] / REFIT_TILE);
}

// the first child of an internal node is the next record (pre-order)
__device__ inline void merge_children(bvh_node* nodes, const int32_t* __restrict__ nright,
                                      int node) {
  const bvh_node A = nodes[node + 1];
  const bvh_node B = nodes[nright[node]];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    nodes[node].lo[d] = fminf(A.lo[d], B.lo[d]);
    nodes[node].hi[d] = fmaxf(A.hi[d], B.hi[d]);
  }
}

__device__ inline uint32_t wg_arrive(uint32_t* p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ void __launch_bounds__(REFIT_THREADS)
    k_refit_tiles(const dbox* __restrict__ ebox, const uint32_t* __restrict__ sorted_ids, int n,
                  bvh_node* nodes, const int32_t* __restrict__ parent,
                  const int32_t* __restrict__ nright, const int32_t* __restrict__ lpre,
                  const int32_t* __restrict__ rmin, const int32_t* __restrict__ rmax,
                  uint32_t* arrive) {
  const int first = blockIdx.x * REFIT_TILE;
  const int last = min(n, first + REFIT_TILE);
  for (int j = first + threadIdx.x; j < last; j += REFIT_THREADS) {
    const int e = (int)sorted_ids[j];
    const dbox b = ebox[e];
    const int leaf = lpre[j];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      nodes[leaf].lo[d] = f32_down(b.lo[d]);
      nodes[leaf].hi[d] = f32_up(b.hi[d]);
    }
    if (n == 1) continue;
    int node = parent[leaf];
    while (node >= 0 && !is_top(rmin, rmax, node)) {
      if (wg_arrive(&arrive[node]) == 0u) break;  // first arrival: sibling not done
      merge_children(nodes, nright, node);
      node = parent[node];
    }
  }
}

// top nodes with no top child: start points of the single-work-group refit
__global__ void k_mark_top(const bvh_node* __restrict__ nodes, const int32_t* __restrict__ nright,
                           const int32_t* __restrict__ rmin, const int32_t* __restrict__ rmax,
                           int n, int32_t* __restrict__ out, int* __restrict__ count) {
  __shared__ int s_cnt, s_base;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool start = false;
  if (i < 2 * n - 1 && nodes[i].left < 0 && is_top(rmin, rmax, i))
    start = !is_top(rmin, rmax, i + 1) && !is_top(rmin, rmax, nright[i]);
  int slot = -1;
  if (start) slot = atomicAdd(&s_cnt, 1);
  __syncthreads();
  if (threadIdx.x == 0) s_base = s_cnt ? atomicAdd(count, s_cnt) : 0;
  __syncthreads();
  if (start) out[s_base + slot] = i;
}

__global__ void __launch_bounds__(1024)
    k_refit_top(bvh_node* nodes, const int32_t* __restrict__ parent,
                const int32_t* __restrict__ nright, const int32_t* __restrict__ rmin,
                const int32_t* __restrict__ rmax, const int32_t* __restrict__ starts, int nstarts,
                uint32_t* arrive) {
  for (int k = threadIdx.x; k < nstarts; k += 1024) {
    int node = starts[k];
    merge_children(nodes, nright, node);
    node = parent[node];
    while (node >= 0) {
      // arrivals needed = number of top children (the others are done)
      const uint32_t need = (uint32_t)(is_top(rmin, rmax, node + 1) ? 1 : 0) +
                            (uint32_t)(is_top(rmin, rmax, nright[node]) ? 1 : 0);
      if (wg_arrive(&arrive[node]) + 1u < need) break;
      merge_children(nodes, nright, node);
      node = parent[node];
    }
  }
}

// (no short circuit, as f64_overlap below)
__device__ inline bool f32_overlap(const bvh_node& a, const float lo[3], const float hi[3]) {
  return (a.lo[0] <= hi[0]) & (a.hi[0] >= lo[0]) & (a.lo[1] <= hi[1]) & (a.hi[1] >= lo[1]) &
         (a.lo[2] <= hi[2]) & (a.hi[2] >= lo[2]);
}

// AABB::isCollid.  All six comparisons are evaluated (no short circuit), so
// the compiler keeps the box loads together in one round trip instead of
// sinking each component's load behind the previous comparison.
__device__ inline bool f64_overlap(const dbox& a, const dbox& b) {
  return (a.lo[0] <= b.hi[0]) & (a.hi[0] >= b.lo[0]) & (a.lo[1] <= b.hi[1]) &
         (a.hi[1] >= b.lo[1]) & (a.lo[2] <= b.hi[2]) & (a.hi[2] >= b.lo[2]);
}

// Reference early rejects for the canonical orientation (a = hi, b = lo):
// 0 kept, REJ_RIGID same rigid surface, REJ_SHARED a shared vertex.
constexpr int REJ_RIGID = 1, REJ_SHARED = 2;
__device__ inline int pair_reject_kind(int4 ta, int4 tb, uint8_t fa) {
  const bool tri_a = ta.z >= 0, tri_b = tb.z >= 0;
  if (tri_a && tri_b) {
    if (ta.w == tb.w && (fa & EF_RIGID)) return REJ_RIGID;  // same rigid surface
    const int va[3] = {ta.x, ta.y, ta.z}, vb[3] = {tb.x, tb.y, tb.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (va[i] == vb[k]) return REJ_SHARED;
    return 0;
  }
  if (!tri_a && !tri_b) {
    const int p[4] = {ta.x, ta.y, tb.x, tb.y};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = i + 1; k < 4; ++k)
        if (p[i] == p[k]) return REJ_SHARED;
    return 0;
  }
  const int4 T = tri_a ? ta : tb, B = tri_a ? tb : ta;
  const int vt[3] = {T.x, T.y, T.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (vt[i] == B.x || vt[i] == B.y) return REJ_SHARED;
  return 0;
}
__device__ inline bool pair_rejected(int4 ta, int4 tb, uint8_t fa) {
  return pair_reject_kind(ta, tb, fa) != 0;
}
// batched scenes: an element's scene is held in the upper half of its
// surface field (the lower half is the surface's id within the scene)
__device__ inline int scene_of(const int4& t) { return t.w >> 16; }

constexpr int TRAV_THREADS = 256;
constexpr int TRAV_BUF = 8;  // pairs buffered per lane before an individual flush

// Persistent full-pass self query.  Lane = one leaf's stackless walk in
// depth-first order after the leaf (every leaf j > i, each subtree once).
// Walk lengths differ a lot within a wavefront, so lanes whose walk has
// ended take the next leaves (a wavefront-aggregated atomic on a per-XCD
// counter) while the others go on; each XCD starts on its own contiguous
// eighth of the leaves (L2 locality) and then helps the others.  Pairs are
// buffered per lane and written by the whole wavefront when a buffer fills.
// DUAL: the boxes are the swept boxes of the first CCD pass; the walk also
// counts the static-box overlaps (the proximity pass's boxes, each inside
// its element's swept box) among the pairs it rejects: a pair with a shared
// vertex always overlaps (both static boxes hold that vertex), a pair on one
// rigid surface is tested.  The kept pairs are tested by k_split_seed, so
// one walk serves both passes.
constexpr int PT_REFILL = 16;   // refill once this many lanes are idle
constexpr int PT_STEPS = 16;    // walk steps between refill checks
//
// Long walks.  A few leaves have boxes that overlap a large part of the
// mesh (a static obstacle's big faces under a cloth): their walks visit tens
// of thousands of nodes one dependent load after another, and the kernel
// waits for the longest.  A walk that reaches TRAV_LONG steps hands the rest
// of its pre-order range [nd, end) to k_traverse_split, where TRAV_SPLIT
// lanes each walk one sub-range.  A rope walk started at any node x visits
// x's subtree and then the following subtrees, i.e. exactly the nodes from x
// on, so walks over [s0, s1) and [s1, end) find the pairs of the walk over
// [s0, end), each once.
constexpr int TRAV_LONG = 512;
constexpr int TRAV_SPLIT = 64;

// The rest of one walk: element e over the pre-order nodes [lo, hi), with the
// FP32 box of its leaf.
struct __align__(16) TravItem {
  int e, lo, hi, pad;
  float blo[4], bhi[4];
};
static_assert(sizeof(TravItem) == TRAV_ITEM_BYTES, "TravItem layout");

template <bool DUAL>
__global__ void __launch_bounds__(TRAV_THREADS) k_traverse(const bvh_node* __restrict__ nodes,
                                                           const int32_t* __restrict__ lpre,
                                                           const dbox* __restrict__ ebox,
                                                           const dbox* __restrict__ pbox,
                                                           const int4* __restrict__ ev,
                                                           const uint8_t* __restrict__ eflags,
                                                           int n, int2* __restrict__ pairs, int cap,
                                                           int* __restrict__ pair_count,
                                                           unsigned long long* __restrict__ cand_count,
                                                           unsigned long long* __restrict__ pcand_count,
                                                           int* __restrict__ next, int scenes,
                                                           TravItem* __restrict__ q, int qcap) {
  __shared__ int2 buf[TRAV_BUF * TRAV_THREADS];
  const int nnodes = 2 * n - 1;
  int* qn = next + kNumXcd;  // items handed to the split kernel
  const int t = threadIdx.x, lane = t & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  // the eighth of the leaves this block's XCD starts on (blocks are dealt
  // to the XCDs round-robin)
  int chunk = (int)(blockIdx.x % kNumXcd), tried = 0;
  auto chunk_end = [&](int c) { return (int)(((int64_t)n * (c + 1)) / kNumXcd); };
  unsigned long long cand = 0, pcand = 0;
  int nbuf = 0;
  int nd = -1;  // current node of this lane's walk, -1: idle
  int e = 0;
  int wlen = 0;  // steps of the current walk
  float mlo[3] = {0.f, 0.f, 0.f}, mhi[3] = {0.f, 0.f, 0.f};
  dbox mb{};
  int4 mt{};
  uint8_t mf = 0;
  auto flush = [&]() {  // the wavefront's buffered pairs
    int incl = nbuf;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int total = __shfl(incl, 63);
    if (total == 0) return;
    int base = 0;
    if (lane == 63) base = atomicAdd(pair_count, total);
    base = __shfl(base, 63) + incl - nbuf;
    for (int k = 0; k < nbuf; ++k)
      if (base + k < cap) pairs[base + k] = buf[k * TRAV_THREADS + t];
    nbuf = 0;
  };
  bool more = n > 1;  // leaves left to take (uniform)
  for (;;) {
    // refill idle lanes
    const uint64_t idle = __ballot(nd < 0);
    if (more && (__popcll(idle) >= PT_REFILL || idle == ~0ull)) {
      const int want = __popcll(idle);
      int got = 0, first = 0;
      while (got == 0 && tried < kNumXcd) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&next[chunk], want);
        base = __shfl(base, 0);
        const int start = (int)(((int64_t)n * chunk) / kNumXcd) + base;
        const int end = chunk_end(chunk);
        if (start < end) {
          got = min(want, end - start);
          first = start;
        } else {
          chunk = (chunk + 1) % kNumXcd;
          ++tried;
        }
      }
      if (got == 0) more = false;
      if (nd < 0) {
        const int r = __popcll(idle & below);
        if (r < got) {
          const int i = first + r;
          const bvh_node me = nodes[lpre[i]];
          e = me.left;
          mb = ebox[e];
          mt = ev[e];
          mf = eflags[e];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            mlo[d] = me.lo[d];
            mhi[d] = me.hi[d];
          }
          nd = me.rope;
          wlen = 0;
        }
      }
    }
    if (!more && __ballot(nd >= 0) == 0) break;
    // walk
    for (int step = 0; step < PT_STEPS && nd >= 0; ++step) {
      if (wlen == TRAV_LONG) {  // hand the rest of this walk over
        const int slot = atomicAdd(qn, 1);
        if (slot < qcap) {
          q[slot] = TravItem{e, nd, nnodes, 0, {mlo[0], mlo[1], mlo[2], 0.f}, {mhi[0], mhi[1], mhi[2], 0.f}};
          nd = -1;
          break;
        }
        // (no room: this lane walks on)
      }
      ++wlen;
      const bvh_node node = nodes[nd];
      if (!f32_overlap(node, mlo, mhi)) {
        nd = node.rope;
        continue;
      }
      if (node.left < 0) {  // internal: its first child is the next record
        ++nd;
        continue;
      }
      nd = node.rope;
      const int e2 = node.left;
      // box, topology and flags in one round trip (the exact overlap test
      // almost always passes where the rounded-out tree box did)
      const dbox ob = ebox[e2];
      const int4 ot = ev[e2];
      const uint8_t of = eflags[e2];
      // (the topology and flags are consumed here so that their loads are
      // issued with the box's, not after the overlap test)
      asm volatile("" ::"v"(ot.x), "v"(ot.y), "v"(ot.z), "v"(ot.w), "v"((int)of));
      if (!f64_overlap(ob, mb)) continue;
      if (scenes && scene_of(ot) != scene_of(mt)) continue;  // batched scenes never pair up
      ++cand;
      const int rk = (e > e2) ? pair_reject_kind(mt, ot, mf) : pair_reject_kind(ot, mt, of);
      if (rk) {
        if constexpr (DUAL)
          if (rk == REJ_SHARED || f64_overlap(pbox[e], pbox[e2])) ++pcand;
        continue;
      }
      buf[nbuf * TRAV_THREADS + t] = make_int2(min(e, e2), max(e, e2));
      if (++nbuf == TRAV_BUF) break;  // the buffer is written before the walk goes on
    }
    if (__ballot(nbuf == TRAV_BUF)) flush();
  }
  flush();
  unsigned long long c = cand;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0 && c) atomicAdd(cand_count, c);
  if constexpr (DUAL) {
    unsigned long long pc = pcand;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o);
    if (lane == 0 && pc) atomicAdd(pcand_count, pc);
  }
}

// Split walks: lane g walks sub-range g % TRAV_SPLIT of item g / TRAV_SPLIT.
// A sub-range walk that again reaches TRAV_LONG steps is handed to the next
// round (qout), if there is one.  Pairs are written with one atomic per
// wavefront and step.
template <bool DUAL>
__global__ void __launch_bounds__(256) k_traverse_split(const bvh_node* __restrict__ nodes,
                                                        const dbox* __restrict__ ebox,
                                                        const dbox* __restrict__ pbox,
                                                        const int4* __restrict__ ev,
                                                        const uint8_t* __restrict__ eflags,
                                                        const TravItem* __restrict__ qin,
                                                        const int* __restrict__ qin_n,
                                                        TravItem* __restrict__ qout,
                                                        int* __restrict__ qout_n, int qcap,
                                                        int2* __restrict__ pairs, int cap,
                                                        int* __restrict__ pair_count,
                                                        unsigned long long* __restrict__ cand_count,
                                                        unsigned long long* __restrict__ pcand_count,
                                                        int scenes) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  const int total = min(*qin_n, qcap) * TRAV_SPLIT;
  unsigned long long cand = 0, pcand = 0;
  const int stride = gridDim.x * blockDim.x;
  for (int g0 = blockIdx.x * blockDim.x; g0 < total; g0 += stride) {  // (uniform per block)
    const int g = g0 + (int)threadIdx.x;
    int nd = -1, s1 = 0, e = 0, wlen = 0;
    float mlo[3] = {0.f, 0.f, 0.f}, mhi[3] = {0.f, 0.f, 0.f};
    dbox mb{};
    int4 mt{};
    uint8_t mf = 0;
    if (g < total) {
      const TravItem it = qin[g / TRAV_SPLIT];
      const int j = g % TRAV_SPLIT;
      const int64_t span = (int64_t)it.hi - it.lo;
      nd = it.lo + (int)(span * j / TRAV_SPLIT);
      s1 = it.lo + (int)(span * (j + 1) / TRAV_SPLIT);
      e = it.e;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mlo[d] = it.blo[d];
        mhi[d] = it.bhi[d];
      }
      mb = ebox[e];
      mt = ev[e];
      mf = eflags[e];
    }
    for (;;) {
      const bool active = nd >= 0 && nd < s1;
      if (__ballot(active) == 0) break;
      bool emit = false;
      int2 pr{};
      if (active) {
        if (wlen == TRAV_LONG && qout) {  // hand the rest of this sub-range over
          const int slot = atomicAdd(qout_n, 1);
          if (slot < qcap) {
            qout[slot] = TravItem{e, nd, s1, 0, {mlo[0], mlo[1], mlo[2], 0.f}, {mhi[0], mhi[1], mhi[2], 0.f}};
            nd = -1;
          }
        }
      }
      if (active && nd >= 0) {
        ++wlen;
        const bvh_node node = nodes[nd];
        if (!f32_overlap(node, mlo, mhi)) {
          nd = node.rope;
        } else if (node.left < 0) {
          ++nd;
        } else {
          nd = node.rope;
          const int e2 = node.left;
          const dbox ob = ebox[e2];
          const int4 ot = ev[e2];
          const uint8_t of = eflags[e2];
          asm volatile("" ::"v"(ot.x), "v"(ot.y), "v"(ot.z), "v"(ot.w), "v"((int)of));
          if (f64_overlap(ob, mb) && !(scenes && scene_of(ot) != scene_of(mt))) {
            ++cand;
            const int rk = (e > e2) ? pair_reject_kind(mt, ot, mf) : pair_reject_kind(ot, mt, of);
            if (rk) {
              if constexpr (DUAL)
                if (rk == REJ_SHARED || f64_overlap(pbox[e], pbox[e2])) ++pcand;
            } else {
              emit = true;
              pr = make_int2(min(e, e2), max(e, e2));
            }
          }
        }
      }
      const uint64_t m = __ballot(emit);
      if (m) {
        int base = 0;
        if (lane == 0) base = atomicAdd(pair_count, (int)__popcll(m));
        base = __shfl(base, 0) + (int)__popcll(m & below);
        if (emit && base < cap) pairs[base] = pr;
      }
    }
  }
  unsigned long long c = cand;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0 && c) atomicAdd(cand_count, c);
  if constexpr (DUAL) {
    unsigned long long pc = pcand;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o);
    if (lane == 0 && pc) atomicAdd(pcand_count, pc);
  }
}

// The swept-box pairs of the dual traversal whose static boxes overlap are
// the proximity pass's stream (and are counted with its candidates).  One
// tile of SPLIT_TILE pairs per work-group, one atomic per tile.
constexpr int SPLIT_ITEMS = 16;
constexpr int SPLIT_TILE = 256 * SPLIT_ITEMS;
__global__ void __launch_bounds__(256) k_split_seed(const int2* __restrict__ seed, int n,
                                                    const dbox* __restrict__ pbox,
                                                    int2* __restrict__ prox,
                                                    int* __restrict__ counters) {
  __shared__ int s_off[256];
  __shared__ int s_base;
  const int t = threadIdx.x;
  const int base_i = blockIdx.x * SPLIT_TILE;
  int2 p[SPLIT_ITEMS];
  uint32_t keep = 0u;
#pragma unroll
  for (int k = 0; k < SPLIT_ITEMS; ++k) {
    const int i = base_i + k * 256 + t;
    p[k] = make_int2(0, 0);
    if (i < n) {
      p[k] = seed[i];
      if (f64_overlap(pbox[p[k].x], pbox[p[k].y])) keep |= 1u << k;
    }
  }
  const int mine = __popc(keep);
  s_off[t] = mine;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the counts
    const int y = t >= off ? s_off[t - off] : 0;
    __syncthreads();
    s_off[t] += y;
    __syncthreads();
  }
  if (t == 255) {
    s_base = s_off[255] ? atomicAdd(&counters[CNT_PPAIRS], s_off[255]) : 0;
    if (s_off[255])
      atomicAdd(reinterpret_cast<unsigned long long*>(counters + CNT_PCAND64),
                (unsigned long long)s_off[255]);
  }
  __syncthreads();
  int o = s_base + s_off[t] - mine;
#pragma unroll
  for (int k = 0; k < SPLIT_ITEMS; ++k)
    if ((keep >> k) & 1u) prox[o++] = p[k];
}

// ---------------------------------------------------------------------------
// Incremental CCD / impact-zone passes.  Between two passes of one step only
// the average velocities of vertices touched by the last apply or zone update
// change ("dirty" vertices).  A pair whose two elements have no dirty vertex
// keeps its swept boxes, its candidate status and its narrow-phase result
// bit for bit, so the next pass carries those over and only
//   * re-queries dirty elements against the refit tree,
//   * re-runs the narrow phase on pairs that touch a dirty element.
// Candidate counts stay exact: count = previous - (pairs touching dirty
// elements under the previous boxes) + (the same under the new boxes).


// dirty_e[e] = element has a dirty vertex; list of dirty elements, each
// once.  The dirty vertices are those the last pass changed: the vertices of
// its hits (vlist, nvl of them) and, after a zone pass, the movable members
// of the impact zones (zlist, nzone of them).  Their elements are found
// through the vertex -> element table, one thread per (vertex, MD_SLOTS-th
// element); a word atomic on the element's flag byte keeps the first finder.
constexpr int MD_SLOTS = 8;
__global__ void __launch_bounds__(256) k_mark_dirty(const int32_t* __restrict__ vlist,
                                                    const int* __restrict__ vcount,
                                                    const uint32_t* __restrict__ zlist, int nzone,
                                                    const uint8_t* __restrict__ fixed,
                                                    const int32_t* __restrict__ vadj_ptr,
                                                    const int32_t* __restrict__ vadj,
                                                    uint8_t* __restrict__ dirty_e,
                                                    int32_t* __restrict__ dlist,
                                                    const int32_t* __restrict__ enode,
                                                    uint8_t* __restrict__ dpre,
                                                    int32_t* __restrict__ rflag,
                                                    int32_t* __restrict__ rlist,
                                                    int* __restrict__ counters) {
  const int nvl = *vcount;
  const long long total = (long long)(nvl + nzone) * MD_SLOTS;
  const long long stride = (long long)gridDim.x * blockDim.x;
  // whole wavefronts iterate together (the appends are per wavefront)
  for (long long t0 = (long long)blockIdx.x * blockDim.x; t0 < total; t0 += stride) {
    const long long t = t0 + threadIdx.x;
    int a = 0, a1 = 0;
    if (t < total) {
      const int k = (int)(t / MD_SLOTS);
      const int v = k < nvl ? vlist[k] : (int)zlist[k - nvl];
      if (k < nvl || !fixed[v]) {
        a = vadj_ptr[v] + (int)(t % MD_SLOTS);
        a1 = vadj_ptr[v + 1];
      }
    }
    for (;; a += MD_SLOTS) {
      const bool live = a < a1;
      if (!__ballot(live)) break;
      int e = 0;
      bool first = false;
      if (live) {
        e = vadj[a];
        const uint32_t bit = 1u << (8 * (e & 3));
        first = !(atomicOr(reinterpret_cast<uint32_t*>(dirty_e + (e & ~3)), bit) & bit);
      }
      const int slot = wave_append(first, &counters[CNT_DIRTY]);
      bool new_range = false;
      int r = 0;
      if (first) {
        dlist[slot] = e;
        // the leaf, and its QW_RANGE-node range of the pre-order layout
        const int p = enode[e];
        dpre[p] = 1;
        r = p >> QW_RANGE_SHIFT;
        new_range = atomicExch(&rflag[r], 1) == 0;
      }
      const int rs = wave_append(new_range, &counters[CNT_DRANGES]);
      if (new_range) rlist[rs] = r;
    }
  }
}

// Incremental refit by growth.  Between full refits only dirty elements
// change their boxes; each dirty leaf gets its exact new FP32 box and then
// grows its ancestors, stopping at the first one that already contains it.
// Internal boxes only grow (they stay valid bounds, possibly loose, until
// the next full refit); growth uses integer atomics on the float bit
// patterns, so concurrent growers of one node compose.  A plain load that
// sees containment is safe to stop on: a node's box only grows during the
// kernel, and whoever grew it carries the growth further up.
__device__ inline void atomic_min_f32(float* a, float v) {
  if (v >= 0.0f) atomicMin(reinterpret_cast<int*>(a), __float_as_int(v));
  else atomicMax(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}
__device__ inline void atomic_max_f32(float* a, float v) {
  if (v >= 0.0f) atomicMax(reinterpret_cast<int*>(a), __float_as_int(v));
  else atomicMin(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}

__global__ void __launch_bounds__(256) k_grow_dirty(const int4* __restrict__ ev,
                                                    const dvec4* __restrict__ x_old,
                                                    const dvec4* __restrict__ av, double dt,
                                                    const int32_t* __restrict__ dlist,
                                                    const int* __restrict__ counters,
                                                    dbox* __restrict__ ebox,
                                                    const int32_t* __restrict__ enode,
                                                    bvh_node* nodes,
                                                    const int32_t* __restrict__ parent, int grow) {
  const int nd = counters[CNT_DIRTY];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nd; k += gridDim.x * blockDim.x) {
    const int e = dlist[k];
    // the element's new swept box
    const int4 t = ev[e];
    const int np = (t.z >= 0) ? 3 : 2;
    const int vid[3] = {t.x, t.y, t.z};
    double blo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
    double bhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int i = 0; i < np; ++i) {
      const dvec4 xo = x_old[vid[i]];
      const dvec4 a = av[vid[i]];
      const double o[3] = {xo.x, xo.y, xo.z};
      const double nn[3] = {xo.x + a.x * dt, xo.y + a.y * dt, xo.z + a.z * dt};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bhi[d] = smax(bhi[d], o[d]);
        bhi[d] = smax(bhi[d], nn[d]);
        blo[d] = smin(blo[d], o[d]);
        blo[d] = smin(blo[d], nn[d]);
      }
    }
    dbox b;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = blo[d] - CS_BOX_EPS;
      b.hi[d] = bhi[d] + CS_BOX_EPS;
    }
    ebox[e] = b;
    if (!grow) continue;
    // the leaf exactly, its ancestors grown
    const int leaf = enode[e];
    float lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = f32_down(b.lo[d]);
      hi[d] = f32_up(b.hi[d]);
      nodes[leaf].lo[d] = lo[d];
      nodes[leaf].hi[d] = hi[d];
    }
    for (int node = parent[leaf]; node >= 0; node = parent[node]) {
      const bvh_node c = nodes[node];
      if ((c.lo[0] <= lo[0]) & (c.lo[1] <= lo[1]) & (c.lo[2] <= lo[2]) & (c.hi[0] >= hi[0]) &
          (c.hi[1] >= hi[1]) & (c.hi[2] >= hi[2]))
        break;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (lo[d] < c.lo[d]) atomic_min_f32(&nodes[node].lo[d], lo[d]);
        if (hi[d] > c.hi[d]) atomic_max_f32(&nodes[node].hi[d], hi[d]);
      }
    }
  }
}

// Query of every dirty element against the whole tree.  A pair of two dirty
// elements is taken from the one earlier in Morton order.  emit = 0: count
// only (previous boxes, into CNT_OLD64); emit = 1: count into CNT_NEW64 and
// append non-rejected pairs to qpairs.
//
// One wavefront per QW_RANGE-node range of the pre-order layout that holds dirty
// leaves: the range's dirty leaves are neighbours in Morton order, so the
// union of their boxes is small.  The wavefront descends the tree breadth
// first with that union box, one level per step with every lane loading one
// node of the frontier, collects the leaves the union box overlaps, and each
// lane then tests its own element against them with the exact boxes.  The
// number of dependent loads is the tree depth rather than the length of a
// single-lane walk.  A frontier or leaf list too large for LDS falls back to
// one stackless walk per lane.
constexpr int QW_FCAP = 1024;  // frontier nodes of one level
constexpr int QW_LCAP = 1024;  // leaves collected by one range

__device__ inline void flush_pairs(int2* buf, int nbuf, int2* __restrict__ qpairs, int qcap,
                                   int* __restrict__ counters) {
  // exclusive scan of nbuf over the wavefront, one atomic per wavefront
  const int lane = threadIdx.x & 63;
  int incl = nbuf;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  const int total = __shfl(incl, 63);
  if (total == 0) return;
  int base = 0;
  if (lane == 63) base = atomicAdd(&counters[CNT_QPAIRS], total);
  base = __shfl(base, 63) + incl - nbuf;
  for (int q = 0; q < nbuf; ++q)
    if (base + q < qcap) qpairs[base + q] = buf[q * 64 + lane];
}

__global__ void __launch_bounds__(64) k_query_ranges(
    const bvh_node* __restrict__ nodes, const int32_t* __restrict__ nright,
    const dbox* __restrict__ ebox, const int4* __restrict__ ev,
    const uint8_t* __restrict__ eflags, const uint8_t* __restrict__ dirty_e,
    const uint8_t* __restrict__ dpre, const int32_t* __restrict__ rlist, int npre, int emit,
    int scenes,
    int2* __restrict__ qpairs, int qcap, int* __restrict__ counters) {
  __shared__ int32_t fr[2][QW_FCAP];
  __shared__ int2 lv[QW_LCAP];  // (pre-order index, element)
  // the range's dirty elements (the queries), compacted: box and pre-order
  // index in one 64 B record (one LDS round trip per query), the rest apart
  struct QueryRec {
    dbox b;
    int32_t si, pad[3];
  };
  __shared__ QueryRec qr[64];
  __shared__ int4 qt[64];
  __shared__ int32_t qe[64];
  __shared__ uint8_t qf[64];
  __shared__ int2 buf[TRAV_BUF * 64];
  const int lane = threadIdx.x;
  const int nr = counters[CNT_DRANGES];
  unsigned long long cand = 0;
  // ranges in block order: the first nr blocks are busy, and the hardware's
  // round-robin placement spreads them over all XCDs
  for (int ri = blockIdx.x; ri < nr; ri += gridDim.x) {
    const int si = QW_RANGE * rlist[ri] + lane;  // this lane's node
    const bool act = lane < QW_RANGE && si < npre && dpre[si];
    bvh_node me{};
    dbox mb{};
    int4 mt{};
    uint8_t mf = 0;
    int e = -1;
    float ulo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, uhi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    if (act) {
      me = nodes[si];
      e = me.left;
      mb = ebox[e];
      mt = ev[e];
      mf = eflags[e];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ulo[d] = me.lo[d];
        uhi[d] = me.hi[d];
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ulo[d] = fminf(ulo[d], __shfl_xor(ulo[d], o));
        uhi[d] = fmaxf(uhi[d], __shfl_xor(uhi[d], o));
      }
    // breadth-first descent with the union box
    if (lane == 0) fr[0][0] = 0;
    __syncthreads();
    int cur = 0, nfr = 1, nlv = 0;
    bool overflow = false;
    const uint64_t below = (1ull << lane) - 1ull;
    while (nfr > 0) {
      int nnext = 0;
      for (int b = 0; b < nfr; b += 64) {
        const int i = b + lane;
        bool kl = false, ki = false;
        int node = 0, c2 = 0, left = -1;
        if (i < nfr) {
          node = fr[cur][i];
          const bvh_node N = nodes[node];
          const int r2 = nright[node];
          if (f32_overlap(N, ulo, uhi)) {
            if (N.left >= 0) {
              kl = true;
              left = N.left;
            } else {
              ki = true;
              c2 = r2;
            }
          }
        }
        const uint64_t ml = __ballot(kl), mi = __ballot(ki);
        if (kl) {
          const int q = nlv + __popcll(ml & below);
          if (q < QW_LCAP) lv[q] = make_int2(node, left);
        }
        if (ki) {
          const int q = nnext + 2 * __popcll(mi & below);
          if (q + 1 < QW_FCAP) {
            fr[cur ^ 1][q] = node + 1;
            fr[cur ^ 1][q + 1] = c2;
          }
        }
        nlv += __popcll(ml);
        nnext += 2 * __popcll(mi);
      }
      __syncthreads();
      if (nlv > QW_LCAP || nnext > QW_FCAP) {
        overflow = true;
        break;
      }
      cur ^= 1;
      nfr = nnext;
    }
    int nbuf = 0;
    // a box-overlapping pair of query element qel and element e2
    auto take = [&](int qel, const int4& qtop, uint8_t qfl, int e2, const int4& ot, uint8_t of) {
      if (scenes && scene_of(ot) != scene_of(qtop)) return;  // batched scenes never pair up
      ++cand;
      if (!emit) return;
      const bool rej = (qel > e2) ? pair_rejected(qtop, ot, qfl) : pair_rejected(ot, qtop, of);
      if (rej) return;
      const int2 pr = make_int2(min(qel, e2), max(qel, e2));
      if (nbuf < TRAV_BUF) {
        buf[nbuf * 64 + lane] = pr;
        ++nbuf;
      } else {
        const int slot = atomicAdd(&counters[CNT_QPAIRS], 1);
        if (slot < qcap) qpairs[slot] = pr;
      }
    };
    if (!overflow) {
      // lanes hold the collected leaves, 64 at a time; the queries are
      // broadcast from LDS
      const uint64_t am = __ballot(act);
      const int na = __popcll(am);
      if (act) {
        const int q = __popcll(am & below);
        qr[q].b = mb;
        qr[q].si = si;
        qt[q] = mt;
        qe[q] = e;
        qf[q] = mf;
      }
      __syncthreads();
      for (int b = 0; b < nlv; b += 64) {
        const bool have = b + lane < nlv;
        int j = -1, e2 = 0;
        dbox ob{};
        int4 ot{};
        uint8_t of = 0, od = 0;
        if (have) {
          const int2 L = lv[b + lane];
          j = L.x;
          e2 = L.y;
          ob = ebox[e2];
          ot = ev[e2];
          of = eflags[e2];
          od = dirty_e[e2];
        }
        // overlaps first (a bit per query), rejects and output afterwards
        uint64_t hits = 0;
#pragma unroll 2
        for (int q = 0; q < na; ++q) {
          const QueryRec Q = qr[q];
          // a pair of two dirty elements is taken from the one earlier in
          // pre-order; the overlap test without branches (AABB::isCollid)
          const bool hit = (j != Q.si) & !(od && j < Q.si) &
                           (ob.lo[0] <= Q.b.hi[0]) & (ob.hi[0] >= Q.b.lo[0]) &
                           (ob.lo[1] <= Q.b.hi[1]) & (ob.hi[1] >= Q.b.lo[1]) &
                           (ob.lo[2] <= Q.b.hi[2]) & (ob.hi[2] >= Q.b.lo[2]);
          hits |= (uint64_t)(have && hit) << q;
        }
        while (hits) {
          const int q = __builtin_ctzll(hits);
          hits &= hits - 1;
          take(qe[q], qt[q], qf[q], e2, ot, of);
        }
      }
    } else if (act) {
      // stackless walk from the root with this lane's leaf box
      int nd = 0;
      while (nd >= 0) {
        const bvh_node node = nodes[nd];
        if (!f32_overlap(node, me.lo, me.hi)) {
          nd = node.rope;
          continue;
        }
        if (node.left < 0) {  // internal: its first child is the next record
          ++nd;
          continue;
        }
        const int j = nd;
        nd = node.rope;
        const int e2 = node.left;
        if (j == si) continue;
        const uint8_t od = dirty_e[e2];
        if (od && j < si) continue;  // counted from e2
        if (!f64_overlap(ebox[e2], mb)) continue;
        take(e, mt, mf, e2, ev[e2], eflags[e2]);
      }
    }
    if (emit) flush_pairs(buf, nbuf, qpairs, qcap, counters);
    __syncthreads();
  }
  unsigned long long c = cand;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0 && c) atomicAdd((unsigned long long*)(counters + (emit ? CNT_NEW64 : CNT_OLD64)), c);
}

// clear the dirty marks of the ranges listed by the previous pass
__global__ void __launch_bounds__(64) k_clear_ranges(const int32_t* __restrict__ rlist, int npre,
                                                     uint8_t* __restrict__ dpre,
                                                     int32_t* __restrict__ rflag,
                                                     const int32_t* __restrict__ dlist,
                                                     uint8_t* __restrict__ dirty_e,
                                                     const int* __restrict__ counters) {
  const int nr = counters[CNT_DRANGES];
  for (int ri = blockIdx.x; ri < nr; ri += gridDim.x) {
    const int r = rlist[ri];
    const int p = QW_RANGE * r + (int)threadIdx.x;
    if ((int)threadIdx.x < QW_RANGE && p < npre) dpre[p] = 0;
    if (threadIdx.x == 0) rflag[r] = 0;
  }
  // the dirty elements' flags
  const int nd = counters[CNT_DIRTY];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nd; k += gridDim.x * blockDim.x)
    dirty_e[dlist[k]] = 0;
}

// carry clean pairs / hits / hit-pair records of the previous pass
constexpr int CARRY_ITEMS = 16;
__global__ void __launch_bounds__(256) k_carry_pairs(const int2* __restrict__ prev, int n,
                                                     const uint8_t* __restrict__ dirty_e,
                                                     int2* __restrict__ cur, int cap,
                                                     int* __restrict__ counters) {
  __shared__ int s_off[256];
  __shared__ int s_base;
  const int t = threadIdx.x;
  const int base_i = blockIdx.x * 256 * CARRY_ITEMS;
  int2 p[CARRY_ITEMS];
  bool keep[CARRY_ITEMS];
  int nk = 0;
#pragma unroll
  for (int q = 0; q < CARRY_ITEMS; ++q) {
    const int i = base_i + q * 256 + t;
    keep[q] = false;
    if (i < n) {
      p[q] = prev[i];
      keep[q] = !dirty_e[p[q].x] && !dirty_e[p[q].y];
    }
    nk += keep[q] ? 1 : 0;
  }
  s_off[t] = nk;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = (t >= off) ? s_off[t - off] : 0;
    __syncthreads();
    s_off[t] += add;
    __syncthreads();
  }
  if (t == 255) s_base = s_off[255] ? atomicAdd(&counters[CNT_PAIRS], s_off[255]) : 0;
  __syncthreads();
  int o = s_base + s_off[t] - nk;
#pragma unroll
  for (int q = 0; q < CARRY_ITEMS; ++q)
    if (keep[q]) {
      if (o < cap) cur[o] = p[q];
      ++o;
    }
}

__global__ void __launch_bounds__(256) k_carry_hits(const hit_rec* __restrict__ prev, int n,
                                                    const uint8_t* __restrict__ dirty_e,
                                                    hit_rec* __restrict__ cur, int cap,
                                                    int* __restrict__ counters) {
  __shared__ int s_cnt, s_base;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool keep = false;
  if (i < n) {
    const uint64_t key = prev[i].key;
    const int lo = (int)(key >> 36), hi = (int)((key >> 8) & 0xFFFFFFFull);
    keep = !dirty_e[lo] && !dirty_e[hi];
  }
  const int slot = block_append(keep, &counters[CNT_HITS], &s_cnt, &s_base);
  if (keep && slot < cap) cur[slot] = prev[i];
}

__global__ void __launch_bounds__(256) k_carry_hp(const int4* __restrict__ prev, int n,
                                                  const uint8_t* __restrict__ dirty_e,
                                                  int4* __restrict__ cur, int cap,
                                                  int* __restrict__ counters) {
  __shared__ int s_cnt, s_base;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int4 r = make_int4(0, 0, 0, 0);
  bool keep = false;
  if (i < n) {
    r = prev[i];
    keep = !dirty_e[r.x] && !dirty_e[r.y];
  }
  const int slot = block_append(keep, &counters[CNT_HIT_PAIRS], &s_cnt, &s_base);
  if (keep) {
    if (slot < cap) cur[slot] = r;
    atomicAdd(&counters[CNT_HIT_SUB], __popc(r.z));
  }
}

void launch_mark_dirty(hipStream_t st, DevState& S, const int* vcount, int nzone) {
  // grid for a typical list; longer lists are walked grid-stride
  hipLaunchKernelGGL(k_mark_dirty, dim3(std::min((MD_SLOTS * S.nv + 255) / 256, 256)), dim3(256),
                     0, st, S.vlist, vcount, S.zlist, nzone, S.fixed, S.vadj_ptr, S.vadj,
                     S.dirty_e, S.dlist, S.leaf_of, S.dpre, S.rflag, S.rlist, S.counters);
}

void launch_clear_ranges(hipStream_t st, DevState& S) {
  const int npre = 2 * S.ne - 1;
  hipLaunchKernelGGL(k_clear_ranges, dim3(std::min((npre + 63) / 64, 2048)), dim3(64), 0, st,
                     S.rlist, npre, S.dpre, S.rflag, S.dlist, S.dirty_e, S.counters);
}

// zero the counters whose bits are set in mask (one launch for several)
__global__ void k_zero_counters(int* __restrict__ counters, uint32_t mask) {
  const int i = threadIdx.x;
  if (i < 32 && ((mask >> i) & 1u)) counters[i] = 0;
}

void launch_zero_counters(hipStream_t st, DevState& S, uint32_t mask) {
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(32), 0, st, S.counters, mask);
}

void launch_grow_dirty(hipStream_t st, DevState& S, double dt) {
  if (S.ne == 0) return;
  // grid for a typical dirty list; longer lists are walked grid-stride
  hipLaunchKernelGGL(k_grow_dirty, dim3(std::min((S.ne + 255) / 256, 512)), dim3(256), 0, st,
                     S.ev, S.x_old, S.av, dt, S.dlist, S.counters, S.ebox, S.leaf_of, S.nodes,
                     S.parent, S.ne >= 2 ? 1 : 0);
}

void launch_query_dirty(hipStream_t st, DevState& S, int emit) {
  if (S.ne < 2) return;
  const int npre = 2 * S.ne - 1;
  hipLaunchKernelGGL(k_query_ranges, dim3(std::min((npre + 63) / 64, S.tune_query)), dim3(64), 0, st,
                     S.nodes, S.nright, S.ebox, S.ev, S.eflags, S.dirty_e, S.dpre, S.rlist, npre,
                     emit, S.nscene > 1, S.qpairs, S.qpair_cap, S.counters);
}

void launch_carry_pairs_from(hipStream_t st, DevState& S, const int2* src, int n) {
  if (n > 0)
    hipLaunchKernelGGL(k_carry_pairs, dim3((n + 256 * CARRY_ITEMS - 1) / (256 * CARRY_ITEMS)), dim3(256), 0, st,
                       src, n, S.dirty_e, S.pairs, S.pair_cap, S.counters);
}

void launch_carry(hipStream_t st, DevState& S, int npairs_prev, int nhits_prev, int nhp_prev) {
  if (npairs_prev > 0)
    hipLaunchKernelGGL(k_carry_pairs, dim3((npairs_prev + 256 * CARRY_ITEMS - 1) / (256 * CARRY_ITEMS)), dim3(256), 0, st,
                       S.pairs_prev, npairs_prev, S.dirty_e, S.pairs, S.pair_cap, S.counters);
  if (nhits_prev > 0)
    hipLaunchKernelGGL(k_carry_hits, dim3((nhits_prev + 255) / 256), dim3(256), 0, st,
                       S.hits_prev, nhits_prev, S.dirty_e, S.hits, S.hit_cap, S.counters);
  if (nhp_prev > 0)
    hipLaunchKernelGGL(k_carry_hp, dim3((nhp_prev + 255) / 256), dim3(256), 0, st, S.hp_prev,
                       nhp_prev, S.dirty_e, S.hp, S.hp_cap, S.counters);
}

// ---------------------------------------------------------------------------
void launch_element_boxes(hipStream_t st, const DevState& S, int moving, double dt, dbox* out) {
  const int ne = S.ne;
  if (ne == 0) return;
  hipLaunchKernelGGL(k_element_boxes, dim3((ne + 255) / 256), dim3(256), 0, st, S.ev, S.x, S.x_old,
                     S.av, dt, moving, ne, out);
}

hipError_t build_bvh(hipStream_t st, DevState& S) {
  const int n = S.ne;
  if (n == 0) return hipSuccess;
  const float init[6] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
  CS_HIP_RET(hipMemcpyAsync(S.bounds, init, sizeof(init), hipMemcpyHostToDevice, st));
  int nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(256), 0, st, S.ebox, n, S.bounds);
  hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, S.ebox, n, S.bounds,
                     S.mcode, S.mid);
  bool in_tmp = false;
  CS_HIP_RET(radix_sort_u32(st, S.mcode, S.mid, S.mcode_tmp, S.mid_tmp, S.sort_hist, S.scan_part, n,
                            30, &in_tmp));
  S.sorted_codes = in_tmp ? S.mcode_tmp : S.mcode;
  S.sorted_ids = in_tmp ? S.mid_tmp : S.mid;
  S.n_top_start = 0;
  if (n == 1) {  // a single leaf: node 0
    bvh_node leaf{};
    leaf.left = 0;
    leaf.rope = -1;
    const int32_t zero = 0, none = -1;
    CS_HIP_RET(hipMemcpyAsync(S.nodes, &leaf, sizeof(leaf), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipMemcpyAsync(S.parent, &none, sizeof(int32_t), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipMemcpyAsync(S.lpre, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipMemcpyAsync(S.leaf_of, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipMemcpyAsync(S.rmin, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipMemcpyAsync(S.rmax, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    CS_HIP_RET(hipStreamSynchronize(st));  // host sources above live on this stack frame
  } else {
    const int nn = 2 * n - 1;
    // every split position is written by k_karras; -1 (end) is a safe default
    CS_HIP_RET(hipMemsetAsync(S.split_right, 0xff, sizeof(int32_t) * (size_t)(n - 1), st));
    hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, st, S.sorted_codes, n,
                       S.kleft, S.kright, S.kparent, S.split_right, S.krmin, S.krmax);
    hipLaunchKernelGGL(k_preorder, dim3((nn + 255) / 256), dim3(256), 0, st, n, S.kleft, S.kparent,
                       S.krmin, S.pre);
    hipLaunchKernelGGL(k_permute, dim3((nn + 255) / 256), dim3(256), 0, st, n, S.pre, S.kright,
                       S.kparent, S.split_right, S.krmin, S.krmax, S.sorted_ids, S.nodes, S.parent,
                       S.nright, S.rmin, S.rmax, S.lpre, S.leaf_of);
    CS_HIP_RET(hipMemsetAsync(S.counters + CNT_TMP, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_mark_top, dim3((nn + 255) / 256), dim3(256), 0, st, S.nodes, S.nright,
                       S.rmin, S.rmax, n, S.top_start, S.counters + CNT_TMP);
    int cnt = 0;
    CS_HIP_RET(hipMemcpyAsync(&cnt, S.counters + CNT_TMP, sizeof(int), hipMemcpyDeviceToHost, st));
    CS_HIP_RET(hipStreamSynchronize(st));
    S.n_top_start = cnt;
  }
  return hipGetLastError();
}

hipError_t refit_bvh(hipStream_t st, DevState& S) {
  const int n = S.ne;
  if (n == 0) return hipSuccess;
  if (n > 1) CS_HIP_RET(hipMemsetAsync(S.arrive, 0, sizeof(uint32_t) * (size_t)(2 * n - 1), st));
  hipLaunchKernelGGL(k_refit_tiles, dim3((n + REFIT_TILE - 1) / REFIT_TILE), dim3(REFIT_THREADS), 0,
                     st, S.ebox, S.sorted_ids, n, S.nodes, S.parent, S.nright, S.lpre, S.rmin,
                     S.rmax, S.arrive);
  if (n > 1 && S.n_top_start > 0)
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, st, S.nodes, S.parent, S.nright,
                       S.rmin, S.rmax, S.top_start, S.n_top_start, S.arrive);
  return hipGetLastError();
}

void launch_traverse(hipStream_t st, DevState& S, int2* out, int cap, const dbox* pbox) {
  const int n = S.ne;
  if (n < 2) return;
  // per-XCD leaf counters, then the item counts of the two split rounds
  (void)hipMemsetAsync(S.tnext, 0, sizeof(int) * (kNumXcd + 2), st);
  const int blocks = std::min(S.num_cus * S.tune_trav, (n + TRAV_THREADS - 1) / TRAV_THREADS);
  TravItem* q0 = reinterpret_cast<TravItem*>(S.tq);
  TravItem* q1 = q0 + TRAV_QCAP;
  int* n0 = S.tnext + kNumXcd;
  int* n1 = n0 + 1;
  auto* cand = (unsigned long long*)(S.counters + CNT_CAND64);
  auto* pcand = pbox ? (unsigned long long*)(S.counters + CNT_PCAND64) : nullptr;
  const int sblocks = 2 * S.num_cus;
  if (pbox) {
    hipLaunchKernelGGL(k_traverse<true>, dim3(std::max(blocks, 1)), dim3(TRAV_THREADS), 0, st,
                       S.nodes, S.lpre, S.ebox, pbox, S.ev, S.eflags, n, out, cap,
                       S.counters + CNT_PAIRS, cand, pcand, S.tnext, S.nscene > 1, q0, TRAV_QCAP);
    hipLaunchKernelGGL(k_traverse_split<true>, dim3(sblocks), dim3(256), 0, st, S.nodes, S.ebox,
                       pbox, S.ev, S.eflags, q0, n0, q1, n1, TRAV_QCAP, out, cap,
                       S.counters + CNT_PAIRS, cand, pcand, S.nscene > 1);
    hipLaunchKernelGGL(k_traverse_split<true>, dim3(sblocks), dim3(256), 0, st, S.nodes, S.ebox,
                       pbox, S.ev, S.eflags, q1, n1, nullptr, nullptr, TRAV_QCAP, out, cap,
                       S.counters + CNT_PAIRS, cand, pcand, S.nscene > 1);
  } else {
    hipLaunchKernelGGL(k_traverse<false>, dim3(std::max(blocks, 1)), dim3(TRAV_THREADS), 0, st,
                       S.nodes, S.lpre, S.ebox, nullptr, S.ev, S.eflags, n, out, cap,
                       S.counters + CNT_PAIRS, cand, nullptr, S.tnext, S.nscene > 1, q0, TRAV_QCAP);
    hipLaunchKernelGGL(k_traverse_split<false>, dim3(sblocks), dim3(256), 0, st, S.nodes, S.ebox,
                       nullptr, S.ev, S.eflags, q0, n0, q1, n1, TRAV_QCAP, out, cap,
                       S.counters + CNT_PAIRS, cand, nullptr, S.nscene > 1);
    hipLaunchKernelGGL(k_traverse_split<false>, dim3(sblocks), dim3(256), 0, st, S.nodes, S.ebox,
                       nullptr, S.ev, S.eflags, q1, n1, nullptr, nullptr, TRAV_QCAP, out, cap,
                       S.counters + CNT_PAIRS, cand, nullptr, S.nscene > 1);
  }
}

void launch_split_seed(hipStream_t st, DevState& S, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_split_seed, dim3((n + SPLIT_TILE - 1) / SPLIT_TILE), dim3(256), 0, st,
                     S.pairs_seed, n, S.pbox, S.pairs, S.counters);
}

}  // namespace cs
