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
//   * k_karras        : radix tree over the sorted codes (index tie-break).
//   * k_refit         : bottom-up FP32 boxes rounded outward, so every FP64
//     overlap is also an FP32 overlap; arrival counters order the updates.
//   * k_traverse      : one lane per leaf in Morton order walks the tree and
//     only visits leaves after itself in that order (each pair once); a
//     leaf overlap is confirmed with the exact FP64 boxes, then the
//     reference's early rejects (shared vertex, same rigid surface;
//     dcollid.cpp:762, 805; dcollid3d.cpp:209-214, 257-264, 279-284) run
//     before the pair is appended to the compacted candidate stream.
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

// Karras (2012) radix tree: internal nodes 0..n-2; child < 0 is leaf ~j.
__global__ void k_karras(const uint32_t* __restrict__ codes, int n, bvh_node* __restrict__ nodes,
                         int32_t* __restrict__ parent, int32_t* __restrict__ lparent,
                         int32_t* __restrict__ rmax) {
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
  const int left = (lo == gamma) ? ~gamma : gamma;
  const int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  nodes[i].left = left;
  nodes[i].right = right;
  rmax[i] = hi;
  if (left < 0) lparent[~left] = i; else parent[left] = i;
  if (right < 0) lparent[~right] = i; else parent[right] = i;
  if (i == 0) parent[0] = -1;
}

// Leaf boxes (sorted order) rounded outward to FP32, then internal nodes
// bottom-up: the second lane to arrive at a node merges both children.
__global__ void k_refit(const dbox* __restrict__ ebox, const uint32_t* __restrict__ sorted_ids,
                        int n, bvh_node* __restrict__ leaves, bvh_node* __restrict__ nodes,
                        const int32_t* __restrict__ parent, const int32_t* __restrict__ lparent,
                        uint32_t* __restrict__ arrive) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int e = (int)sorted_ids[j];
  const dbox b = ebox[e];
  bvh_node lf;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lf.lo[d] = f32_down(b.lo[d]);
    lf.hi[d] = f32_up(b.hi[d]);
  }
  lf.left = e;
  lf.right = j;
  leaves[j] = lf;
  if (n == 1) return;
  __threadfence();
  int node = lparent[j];
  while (node >= 0) {
    if (atomicAdd(&arrive[node], 1u) == 0u) return;  // first arrival stops
    __threadfence();
    const int l = nodes[node].left, r = nodes[node].right;
    const bvh_node A = (l < 0) ? leaves[~l] : nodes[l];
    const bvh_node B = (r < 0) ? leaves[~r] : nodes[r];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      nodes[node].lo[d] = fminf(A.lo[d], B.lo[d]);
      nodes[node].hi[d] = fmaxf(A.hi[d], B.hi[d]);
    }
    __threadfence();
    node = parent[node];
  }
}

__device__ inline bool f32_overlap(const bvh_node& a, const float lo[3], const float hi[3]) {
  return (a.lo[0] <= hi[0] && a.hi[0] >= lo[0]) && (a.lo[1] <= hi[1] && a.hi[1] >= lo[1]) &&
         (a.lo[2] <= hi[2] && a.hi[2] >= lo[2]);
}

__device__ inline bool f64_overlap(const dbox& a, const dbox& b) {  // AABB::isCollid
  return (a.lo[0] <= b.hi[0] && a.hi[0] >= b.lo[0]) && (a.lo[1] <= b.hi[1] && a.hi[1] >= b.lo[1]) &&
         (a.lo[2] <= b.hi[2] && a.hi[2] >= b.lo[2]);
}

// Reference early rejects for the canonical orientation (a = hi, b = lo).
__device__ inline bool pair_rejected(int4 ta, int4 tb, uint8_t fa) {
  const bool tri_a = ta.z >= 0, tri_b = tb.z >= 0;
  if (tri_a && tri_b) {
    if (ta.w == tb.w && (fa & EF_RIGID)) return true;  // same rigid surface
    const int va[3] = {ta.x, ta.y, ta.z}, vb[3] = {tb.x, tb.y, tb.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (va[i] == vb[k]) return true;
    return false;
  }
  if (!tri_a && !tri_b) {
    const int p[4] = {ta.x, ta.y, tb.x, tb.y};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = i + 1; k < 4; ++k)
        if (p[i] == p[k]) return true;
    return false;
  }
  const int4 T = tri_a ? ta : tb, B = tri_a ? tb : ta;
  const int vt[3] = {T.x, T.y, T.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (vt[i] == B.x || vt[i] == B.y) return true;
  return false;
}

constexpr int TRAV_STACK = 64;

__global__ void __launch_bounds__(256) k_traverse(const bvh_node* __restrict__ nodes,
                                                  const bvh_node* __restrict__ leaves,
                                                  const int32_t* __restrict__ rmax,
                                                  const dbox* __restrict__ ebox,
                                                  const int4* __restrict__ ev,
                                                  const uint8_t* __restrict__ eflags, int n,
                                                  int2* __restrict__ pairs, int cap,
                                                  int* __restrict__ pair_count,
                                                  unsigned long long* __restrict__ cand_count,
                                                  int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cand = 0;
  if (i < n && n > 1) {
    const bvh_node me = leaves[i];
    const int e = me.left;
    const dbox mb = ebox[e];
    const int4 mt = ev[e];
    const uint8_t mf = eflags[e];
    int stack[TRAV_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int nd = stack[--sp];
      const bvh_node node = nodes[nd];
      const int ch[2] = {node.left, node.right};
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int cc = ch[c];
        if (cc < 0) {
          const int j = ~cc;
          if (j <= i) continue;
          const bvh_node lf = leaves[j];
          if (!f32_overlap(lf, me.lo, me.hi)) continue;
          const int e2 = lf.left;
          const dbox ob = ebox[e2];
          if (!f64_overlap(ob, mb)) continue;
          ++cand;
          const int4 ot = ev[e2];
          const int hi = max(e, e2), lo = min(e, e2);
          const bool rej = (e > e2) ? pair_rejected(mt, ot, mf) : pair_rejected(ot, mt, eflags[e2]);
          if (rej) continue;
          const int slot = atomicAdd(pair_count, 1);
          if (slot < cap) pairs[slot] = make_int2(lo, hi);
        } else {
          if (rmax[cc] <= i) continue;
          if (!f32_overlap(nodes[cc], me.lo, me.hi)) continue;
          if (sp < TRAV_STACK) {
            stack[sp++] = cc;
          } else {
            atomicOr(err, CS_DEVERR_STACK);
          }
        }
      }
    }
  }
  // one 64-bit atomic per wavefront for the candidate count
  for (int off = 32; off > 0; off >>= 1) cand += __shfl_down(cand, off);
  if ((threadIdx.x & 63) == 0 && cand) atomicAdd(cand_count, cand);
}

// ---------------------------------------------------------------------------
void launch_element_boxes(hipStream_t st, const DevState& S, int moving, double dt) {
  const int ne = S.ne;
  if (ne == 0) return;
  hipLaunchKernelGGL(k_element_boxes, dim3((ne + 255) / 256), dim3(256), 0, st, S.ev, S.x, S.x_old,
                     S.av, dt, moving, ne, S.ebox);
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
  CS_HIP_RET(radix_sort_u32(st, S.mcode, S.mid, S.mcode_tmp, S.mid_tmp, S.sort_hist, n, 30, &in_tmp));
  S.sorted_codes = in_tmp ? S.mcode_tmp : S.mcode;
  S.sorted_ids = in_tmp ? S.mid_tmp : S.mid;
  if (n > 1)
    hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, st, S.sorted_codes, n,
                       S.nodes, S.parent, S.lparent, S.rmax);
  return hipGetLastError();
}

hipError_t refit_bvh(hipStream_t st, DevState& S) {
  const int n = S.ne;
  if (n == 0) return hipSuccess;
  if (n > 1) CS_HIP_RET(hipMemsetAsync(S.arrive, 0, sizeof(uint32_t) * (size_t)(n - 1), st));
  hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, st, S.ebox, S.sorted_ids, n,
                     S.leaves, S.nodes, S.parent, S.lparent, S.arrive);
  return hipGetLastError();
}

void launch_traverse(hipStream_t st, DevState& S, int cap) {
  const int n = S.ne;
  if (n == 0) return;
  hipLaunchKernelGGL(k_traverse, dim3((n + 255) / 256), dim3(256), 0, st, S.nodes, S.leaves,
                     S.rmax, S.ebox, S.ev, S.eflags, n, S.pairs, cap, S.counters + CNT_PAIRS,
                     S.cand_count, S.err);
}

}  // namespace cs
