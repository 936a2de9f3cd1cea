// Response and per-vertex stages for gfx950.
//
//  * apply_hits: hit records are sorted by their canonical key (pair (lo,hi),
//    sub-test), expanded to one record per deformable point in that order,
//    stably sorted by vertex, and each vertex segment is summed left to right
//    from 0.0 by one lane -- the same sequence of additions the reference
//    performs into collsnImpulse / friction (dcollid3d.cpp:1053-1079,
//    1235-1284).  The Jacobi update avgVel += (imp + fric) / count is
//    updateAverageVelocity (dcollid.cpp:677-751).  No atomics touch values,
//    so results are deterministic.
//  * update_zones: union-find roots (smallest vertex id of each zone) ->
//    stable sort of vertices by root -> one lane per zone computes the rigid
//    motion of updateImpactListVelocity (dcollid3d.cpp:70-200) over its
//    members in ascending vertex id.
//  * strain_limit: reduceSuperelast (dcollid.cpp:486-596) is an in-place
//    Gauss-Seidel sweep in element order.  Elements are bucketed into levels
//    (level = 1 + max level of earlier elements sharing a vertex); elements of
//    one level share no vertex and each vertex sees its elements in the
//    original order, so a level-by-level sweep reproduces the sequential
//    one exactly.  A parallel pre-check skips sweeps in which no edge is over
//    the limit (such a sweep changes nothing and ends the loop).
#include <algorithm>

#include "cs_internal.h"

namespace cs {

__global__ void k_predict(dvec4* __restrict__ x, dvec4* __restrict__ x_old,
                          const dvec4* __restrict__ vel, double dt, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const dvec4 p = x[v], u = vel[v];
  dvec4 o = p;
  dvec4 n;
  n.x = o.x + dt * u.x;
  n.y = o.y + dt * u.y;
  n.z = o.z + dt * u.z;
  n.w = 0.0;
  x_old[v] = o;
  x[v] = n;
}

// computeAverageVelocity (dcollid.cpp:160-220)
__global__ void k_average_velocity(dvec4* __restrict__ x, const dvec4* __restrict__ x_old,
                                   dvec4* __restrict__ av, double dt, int nv,
                                   int* __restrict__ err) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const dvec4 p = x[v], o = x_old[v];
  dvec4 a;
  if (dt > CS_ROUND_EPS) {
    a.x = (p.x - o.x) / dt;
    a.y = (p.y - o.y) / dt;
    a.z = (p.z - o.z) / dt;
  } else {
    a.x = a.y = a.z = 0.0;
  }
  a.w = 0.0;
  if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) atomicOr(err, CS_DEVERR_NONFINITE);
  av[v] = a;
  x[v] = o;
}

__global__ void k_uf_init(int32_t* __restrict__ uf, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) uf[v] = v;
}

__global__ void k_hit_keys(const hit_rec* __restrict__ hits, int n, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = hits[i].key;
  idx[i] = (uint32_t)i;
}

// one record per (hit in canonical order, point); static points get the
// sentinel key nv and sort last
__global__ void k_expand_records(const hit_rec* __restrict__ hits,
                                 const uint32_t* __restrict__ sorted_idx, int n, int nv,
                                 uint32_t* __restrict__ rkey, uint32_t* __restrict__ rval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * n) return;
  const int r = i >> 2, p = i & 3;
  const uint32_t h = sorted_idx[r];
  const int v = hits[h].v[p];
  rkey[i] = (v >= 0) ? (uint32_t)v : (uint32_t)nv;
  rval[i] = h * 4u + (uint32_t)p;
}

__global__ void k_apply(const uint32_t* __restrict__ rkey, const uint32_t* __restrict__ rval,
                        int nrec, int nv, const hit_rec* __restrict__ hits,
                        dvec4* __restrict__ av, uint8_t* __restrict__ has,
                        uint8_t* __restrict__ dirty_v, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  const uint32_t v = rkey[i];
  if (v >= (uint32_t)nv) return;
  if (i > 0 && rkey[i - 1] == v) return;  // not the head of the vertex segment
  double imp[3] = {0.0, 0.0, 0.0}, fric[3] = {0.0, 0.0, 0.0};
  int count = 0;
  for (int j = i; j < nrec && rkey[j] == v; ++j) {
    const uint32_t hv = rval[j];
    const hit_rec& h = hits[hv >> 2];
    const int p = hv & 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      imp[k] += h.imp[p][k];
      fric[k] += h.fric[p][k];
    }
    ++count;
  }
  dvec4 a = av[v];
  a.x += (imp[0] + fric[0]) / count;
  a.y += (imp[1] + fric[1]) / count;
  a.z += (imp[2] + fric[2]) / count;
  if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) atomicOr(err, CS_DEVERR_NONFINITE);
  av[v] = a;
  has[v] = 1;
  dirty_v[v] = 1;
}

static int bits_for(uint64_t maxval) {  // bits needed for values <= maxval, rounded to 8
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  b = (b + 7) & ~7;
  return b < 8 ? 8 : b;
}

hipError_t apply_hits(hipStream_t st, DevState& S, int nhits) {
  if (nhits <= 0) return hipSuccess;
  const int nb = (nhits + 255) / 256;
  hipLaunchKernelGGL(k_hit_keys, dim3(nb), dim3(256), 0, st, S.hits, nhits, S.hkey, S.hidx);
  bool in_tmp = false;
  const int kbits = bits_for(hit_key((uint32_t)S.ne, (uint32_t)S.ne, 255u));
  CS_HIP_RET(radix_sort_u64(st, S.hkey, S.hidx, S.hkey_tmp, S.hidx_tmp, S.sort_hist, S.scan_part, nhits, kbits,
                            &in_tmp));
  const uint32_t* sidx = in_tmp ? S.hidx_tmp : S.hidx;
  const int nrec = 4 * nhits;
  hipLaunchKernelGGL(k_expand_records, dim3((nrec + 255) / 256), dim3(256), 0, st, S.hits, sidx,
                     nhits, S.nv, S.rkey, S.rval);
  CS_HIP_RET(radix_sort_u32(st, S.rkey, S.rval, S.rkey_tmp, S.rval_tmp, S.sort_hist, S.scan_part, nrec,
                            bits_for((uint64_t)S.nv), &in_tmp));
  const uint32_t* rk = in_tmp ? S.rkey_tmp : S.rkey;
  const uint32_t* rv = in_tmp ? S.rval_tmp : S.rval;
  hipLaunchKernelGGL(k_apply, dim3((nrec + 255) / 256), dim3(256), 0, st, rk, rv, nrec, S.nv,
                     S.hits, S.av, S.has, S.dirty_v, S.counters + CNT_ERR);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// impact zones
__device__ inline int uf_root(const int32_t* uf, int x) {
  int p = uf[x];
  while (p != x) {
    x = p;
    p = uf[x];
  }
  return x;
}

// sort keys of the zone members: (root << vbits) | vertex, so one sort
// orders zones by root and members by ascending vertex id
__global__ void k_zone_keys(const uint32_t* __restrict__ zlist, int n, const int32_t* __restrict__ uf,
                            int vbits, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = zlist[i];
  const uint32_t r = (uint32_t)uf_root(uf, (int)v);
  key[i] = ((uint64_t)r << vbits) | v;
  val[i] = v;
}

// per sorted member: its root; zone heads (first sorted index) and ends (by root)
__global__ void __launch_bounds__(256) k_zone_heads(const uint64_t* __restrict__ key, int n,
                                                   int vbits, uint32_t* __restrict__ zroot,
                                                   int32_t* __restrict__ zhead,
                                                   int32_t* __restrict__ zend,
                                                   int* __restrict__ counters) {
  __shared__ int s_cnt, s_base;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool head = false;
  if (i < n) {
    const uint32_t z = (uint32_t)(key[i] >> vbits);
    zroot[i] = z;
    head = (i == 0) || (uint32_t)(key[i - 1] >> vbits) != z;
    if (i == n - 1 || (uint32_t)(key[i + 1] >> vbits) != z) zend[z] = i + 1;
  }
  const int slot = block_append(head, &counters[CNT_ZONES], &s_cnt, &s_base);
  if (head) zhead[slot] = i;
}

// updateImpactListVelocity (dcollid3d.cpp:70-200), one 64-lane wavefront per
// zone.  Sums run in the canonical lane order: lane l adds the terms of
// members l, l+64, ... (ascending vertex id) from 0.0, then the 64 lane sums
// are added in lane order.
constexpr int ZONE_LANES = 64;

__device__ inline void lane_total(double (*part)[9], int nout, double* tot) {
  for (int k = 0; k < nout; ++k) tot[k] = 0.0;
  for (int l = 0; l < ZONE_LANES; ++l)
    for (int k = 0; k < nout; ++k) tot[k] = tot[k] + part[l][k];
}

__global__ void __launch_bounds__(ZONE_LANES)
    k_zone_update(const uint32_t* __restrict__ zkey, const uint32_t* __restrict__ zval,
                  const int32_t* __restrict__ zhead, const int32_t* __restrict__ zend,
                  const dvec4* __restrict__ x_old, dvec4* __restrict__ av,
                  const uint8_t* __restrict__ fixed, double m, double dt,
                  int* __restrict__ counters, uint8_t* __restrict__ dirty_v) {
  __shared__ double part[ZONE_LANES][9];
  int* err = counters + CNT_ERR;
  const int l = threadIdx.x;
  const int nz = counters[CNT_ZONES];
  for (int z = blockIdx.x; z < nz; z += gridDim.x) {
    const int b = zhead[z];
    const int e = zend[zkey[b]];
    const int n = e - b;
    // x_cm, v_cm
    double t[9];
    for (int k = 0; k < 6; ++k) t[k] = 0.0;
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const int p = zval[j];
      const dvec4 o = x_old[p], a = av[p];
      t[0] = o.x + t[0]; t[1] = o.y + t[1]; t[2] = o.z + t[2];
      t[3] = a.x + t[3]; t[4] = a.y + t[4]; t[5] = a.z + t[5];
    }
    for (int k = 0; k < 6; ++k) part[l][k] = t[k];
    __syncthreads();
    double tot[9];
    lane_total(part, 6, tot);
    const double x_cm[3] = {tot[0] / n, tot[1] / n, tot[2] / n};
    const double v_cm[3] = {tot[3] / n, tot[4] / n, tot[5] / n};
    __syncthreads();
    // angular momentum L = sum m dx x dv
    for (int k = 0; k < 3; ++k) t[k] = 0.0;
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const int p = zval[j];
      const dvec4 o = x_old[p], a = av[p];
      const double xo[3] = {o.x, o.y, o.z}, vv[3] = {a.x, a.y, a.z};
      double dx[3], dv[3], Li[3];
      sub3(xo, x_cm, dx);
      sub3(vv, v_cm, dv);
      cross3(dx, dv, Li);
      for (int k = 0; k < 3; ++k) t[k] = m * Li[k] + t[k];
    }
    for (int k = 0; k < 3; ++k) part[l][k] = t[k];
    __syncthreads();
    double L[3];
    lane_total(part, 3, L);
    __syncthreads();
    // inertia tensor I = sum m (|dx|^2 delta - dx dx^T)
    for (int k = 0; k < 9; ++k) t[k] = 0.0;
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const int p = zval[j];
      const dvec4 o = x_old[p];
      const double xo[3] = {o.x, o.y, o.z};
      double dx[3];
      sub3(xo, x_cm, dx);
      const double mag_dx = mag3(dx);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double tm = -dx[r] * dx[c];
          if (r == c) tm += mag_dx * mag_dx;
          t[3 * r + c] = tm * m + t[3 * r + c];
        }
    }
    for (int k = 0; k < 9; ++k) part[l][k] = t[k];
    __syncthreads();
    double Iv[9];
    lane_total(part, 9, Iv);
    __syncthreads();
    double I[3][3], tmp[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) I[r][c] = Iv[3 * r + c];
    double w[3];
    for (int c = 0; c < 3; ++c) {
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) tmp[r][q] = I[r][q];
      for (int r = 0; r < 3; ++r) tmp[r][c] = L[r];
      if (det3(I) < CS_ROUND_EPS) w[c] = 0.0;
      else w[c] = det3(tmp) / det3(I);
    }
    const double mag_w = mag3(w);
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const int p = zval[j];
      if (fixed[p]) continue;
      const dvec4 o = x_old[p];
      const double xo[3] = {o.x, o.y, o.z};
      double dx[3], xF[3], xR[3], wxR[3], tmpV[3], x_new[3], an[3];
      sub3(xo, x_cm, dx);
      if (mag_w < CS_ROUND_EPS) {
        for (int k = 0; k < 3; ++k) {
          xF[k] = dx[k];
          wxR[k] = 0.0;
        }
        sub3(dx, xF, xR);
      } else {
        const double sc = dot3(dx, w) / dot3(w, w);
        for (int k = 0; k < 3; ++k) xF[k] = sc * w[k];
        sub3(dx, xF, xR);
        const double sw = cs_det_sin(dt * mag_w) / mag_w;
        for (int k = 0; k < 3; ++k) tmpV[k] = sw * w[k];
        cross3(tmpV, xR, wxR);
      }
      for (int k = 0; k < 3; ++k) {
        x_new[k] = x_cm[k] + dt * v_cm[k] + xF[k] + cs_det_cos(dt * mag_w) * xR[k] + wxR[k];
        an[k] = (x_new[k] - xo[k]) / dt;
      }
      if (isnan(an[0]) || isnan(an[1]) || isnan(an[2])) atomicOr(err, CS_DEVERR_NONFINITE);
      av[p] = dvec4{an[0], an[1], an[2], 0.0};
      dirty_v[p] = 1;
    }
  }
}

// Every vertex of a zone with two or more members took part in a union, so
// the union calls list the members as they go (zone_touch); only the listed
// vertices are sorted, not the whole mesh.
hipError_t update_zones(hipStream_t st, DevState& S, double m, double dt, int nlist) {
  if (nlist <= 0) return hipSuccess;
  const int nb = (nlist + 255) / 256;
  int vbits = 1;
  while (vbits < 32 && ((uint64_t)(S.nv - 1) >> vbits) != 0) ++vbits;
  hipLaunchKernelGGL(k_zone_keys, dim3(nb), dim3(256), 0, st, S.zlist, nlist, S.uf, vbits,
                     S.zkey64, S.zval);
  bool in_tmp = false;
  CS_HIP_RET(radix_sort_u64(st, S.zkey64, S.zval, S.zkey64_tmp, S.zval_tmp, S.sort_hist,
                            S.scan_part, nlist, bits_for(((uint64_t)(S.nv - 1) << vbits) | (uint64_t)(S.nv - 1)),
                            &in_tmp));
  const uint64_t* zk = in_tmp ? S.zkey64_tmp : S.zkey64;
  const uint32_t* zv = in_tmp ? S.zval_tmp : S.zval;
  hipLaunchKernelGGL(k_zone_heads, dim3(nb), dim3(256), 0, st, zk, nlist, vbits, S.zkey, S.zhead,
                     S.zend, S.counters);
  hipLaunchKernelGGL(k_zone_update, dim3(std::min(nlist, 2048)), dim3(ZONE_LANES), 0, st, S.zkey,
                     zv, S.zhead, S.zend, S.x_old, S.av, S.fixed, m, dt, S.counters, S.dirty_v);
  return hipGetLastError();
}

__global__ void k_zone_labels(const int32_t* __restrict__ uf, int nv, int32_t* __restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) out[v] = uf_root(uf, v);
}

void launch_zone_labels(hipStream_t st, DevState& S, int32_t* out) {
  hipLaunchKernelGGL(k_zone_labels, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.uf, S.nv, out);
}

// ---------------------------------------------------------------------------
// detectDomainBoundaryCollision (dcollid.cpp:116-158); a point's second visit
// is a no-op, so each point is handled once
__global__ void k_domain(dvec4* __restrict__ av, dvec4* __restrict__ x,
                         const dvec4* __restrict__ x_old, uint8_t* __restrict__ has, int nv,
                         double lo0, double lo1, double lo2, double hi0, double hi1, double hi2,
                         double dt, double mu, double thickness) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  dvec4 a4 = av[v], o4 = x_old[v], c4 = x[v];
  double a[3] = {a4.x, a4.y, a4.z}, o[3] = {o4.x, o4.y, o4.z}, c[3] = {c4.x, c4.y, c4.z};
  const double L[3] = {lo0, lo1, lo2}, U[3] = {hi0, hi1, hi2};
  double dv = 0;
  bool touched = false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double cand = o[j] + dt * a[j];
    if (cand <= L[j]) {
      touched = true;
      dv = fabs(a[j]);
      a[j] = 0.0;
      c[j] = L[j] + thickness;
    } else if (cand >= U[j]) {
      touched = true;
      dv = fabs(a[j]);
      a[j] = 0.0;
      c[j] = U[j] - thickness;
    }
  }
  const double preVt = mag3(a);
  if (preVt > CS_MACH_EPS) {
    const double f = smax(1.0 - mu * dv / preVt, 0.0);
#pragma unroll
    for (int j = 0; j < 3; ++j) a[j] *= f;
  }
  if (touched) has[v] = 1;
  av[v] = dvec4{a[0], a[1], a[2], 0.0};
  x[v] = dvec4{c[0], c[1], c[2], 0.0};
}

void launch_domain_boundary(hipStream_t st, DevState& S, const double lo[3], const double hi[3],
                            double dt, double mu, double thickness) {
  hipLaunchKernelGGL(k_domain, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.av, S.x, S.x_old,
                     S.has, S.nv, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], dt, mu, thickness);
}

// updateFinalPosition (dcollid.cpp:562-584)
__global__ void k_final_position(dvec4* __restrict__ x, const dvec4* __restrict__ x_old,
                                 const dvec4* __restrict__ av, double dt, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const dvec4 o = x_old[v], a = av[v];
  x[v] = dvec4{o.x + a.x * dt, o.y + a.y * dt, o.z + a.z * dt, 0.0};
}

void launch_final_position(hipStream_t st, DevState& S, double dt) {
  hipLaunchKernelGGL(k_final_position, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.x, S.x_old,
                     S.av, dt, S.nv);
}

// updateFinalVelocity (dcollid.cpp:598-624)
__global__ void k_final_velocity(dvec4* __restrict__ vel, const dvec4* __restrict__ av,
                                 const uint8_t* __restrict__ has, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (has[v]) {
    const dvec4 a = av[v];
    vel[v] = dvec4{a.x, a.y, a.z, 0.0};
  }
}

void launch_final_velocity(hipStream_t st, DevState& S) {
  hipLaunchKernelGGL(k_final_velocity, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.vel, S.av,
                     S.has, S.nv);
}

void launch_predict(hipStream_t st, DevState& S, double dt) {
  hipLaunchKernelGGL(k_predict, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.x, S.x_old, S.vel,
                     dt, S.nv);
}

void launch_average_velocity(hipStream_t st, DevState& S, double dt) {
  hipLaunchKernelGGL(k_average_velocity, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.x, S.x_old,
                     S.av, dt, S.nv, S.counters + CNT_ERR);
}

void launch_uf_init(hipStream_t st, DevState& S) {
  hipLaunchKernelGGL(k_uf_init, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.uf, S.nv);
}

// ---------------------------------------------------------------------------
// strain limiting
// edge test with the x_old length given (it is fixed during the step)
__device__ __attribute__((always_inline)) inline bool edge_over_limit_l(const double* o0, const double* o1, const dvec4& a0,
                                         const dvec4& a1, double len_old, double len0, double dt,
                                         bool* degenerate) {
  const double xc0[3] = {o0[0] + dt * a0.x, o0[1] + dt * a0.y, o0[2] + dt * a0.z};
  const double xc1[3] = {o1[0] + dt * a1.x, o1[1] + dt * a1.y, o1[2] + dt * a1.z};
  double sq = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double d = xc0[i] - xc1[i];
    sq += d * d;
  }
  *degenerate = false;
  // Fast decision, taken only far from every threshold: the length and the
  // two quotients from the hardware reciprocal / reciprocal square root
  // (accurate to far better than 1e-6 relative) differ from their correctly
  // rounded values by much less than the 1e-4 margin, so the outcome equals
  // the exact test's.  Inputs near a threshold, degenerate or non-finite
  // ones take the exact path below.
  if (sq > 0.0 && len_old > 2.0 * CS_ROUND_EPS) {
    const double ln = sq * __builtin_amdgcn_rsq(sq);
    if (ln > 2.0 * CS_ROUND_EPS) {
      const double rate = (ln - len_old) * __builtin_amdgcn_rcp(len_old);
      const double strn = (ln - len0) * __builtin_amdgcn_rcp(len0);
      const double m = fmax(fabs(rate), fabs(strn));
      if (m > CS_SUPERELAS_TOL + 1e-4) return true;
      if (m < CS_SUPERELAS_TOL - 1e-4) return false;
    }
  }
  const double len_new = sqrt(sq);  // dist3 of the same points
  if (len_old > CS_ROUND_EPS && len_new > CS_ROUND_EPS) {
    const double strain_rate = (len_new - len_old) / len_old;
    const double strain = (len_new - len0) / len0;
    return fabs(strain) > CS_SUPERELAS_TOL || fabs(strain_rate) > CS_SUPERELAS_TOL;
  }
  *degenerate = true;
  return true;
}

// Sparse exact sweep.  An element none of whose vertices was written earlier
// in the sweep reads the pre-sweep velocities, so whether it writes is
// decided by the parallel pre-check.  Only pre-check-active elements and
// later visits of written ("dirty") vertices are evaluated, in level order.
//
// Dirty state travels along each vertex's chain of visits: a slot that
// evaluates a dirty vertex marks that vertex in the slot of its next visit,
// so every later element holding the vertex is evaluated, as in a
// sequential sweep.  Almost every next visit is on the following level; such
// marks go to an LDS table for that level, the rest to the slot's flag word
// in memory (bytes 0-2 dirty vertices, byte 3 pre-check active) with an LDS
// bit telling that level to re-read its flag words.  The heads and flag
// words of the following level and of the next pre-check-active level are
// loaded one level ahead, so a level's critical path is: LDS marks -> body and
// velocities of the evaluated slots (one memory round trip) -> arithmetic ->
// barrier.  Levels with nothing to do are skipped using LDS level bits.

// x_old of the slot vertices and edge lengths, once per step
__global__ void k_strain_pack(const StrainHead* __restrict__ head, StrainBody* __restrict__ body,
                              const dvec4* __restrict__ x_old, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const StrainHead h = head[s];
  StrainBody& r = body[s];
  const dvec4 a = x_old[h.v[0]], b = x_old[h.v[1]];
  const double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
  r.x[0] = a.x; r.x[1] = a.y; r.x[2] = a.z;
  r.x[3] = b.x; r.x[4] = b.y; r.x[5] = b.z;
  r.lold[0] = dist3(pa, pb);
  if (h.v[2] >= 0) {
    const dvec4 c = x_old[h.v[2]];
    const double pc[3] = {c.x, c.y, c.z};
    r.x[6] = c.x; r.x[7] = c.y; r.x[8] = c.z;
    r.lold[1] = dist3(pb, pc);
    r.lold[2] = dist3(pc, pa);
  }
}

// pre-check with the current velocities: flag words (dirty bytes cleared,
// active byte set), lvl_active and the any-edge flag
__global__ void k_strain_check(const StrainHead* __restrict__ head,
                               const StrainBody* __restrict__ body, const dvec4* __restrict__ av,
                               double dt, int n, uint32_t* __restrict__ sflag,
                               uint8_t* __restrict__ lvl_active, int* __restrict__ counters) {
  if (counters[CNT_STRAIN_FLAG] < 0) return;  // loop already finished
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  bool any = false;
  if (s < n) {
    const StrainHead h = head[s];
    const StrainBody& r = body[s];
    const dvec4 a0 = av[h.v[0]], a1 = av[h.v[1]];
    bool deg;
    if (edge_over_limit_l(r.x, r.x + 3, a0, a1, r.lold[0], r.L0[0], dt, &deg)) any = true;
    if (h.v[2] >= 0) {
      const dvec4 a2 = av[h.v[2]];
      if (edge_over_limit_l(r.x + 3, r.x + 6, a1, a2, r.lold[1], r.L0[1], dt, &deg)) any = true;
      if (edge_over_limit_l(r.x + 6, r.x, a2, a0, r.lold[2], r.L0[2], dt, &deg)) any = true;
    }
    sflag[s] = any ? 0x01000000u : 0u;
    if (any) lvl_active[h.level] = 1;
  }
  if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(&counters[CNT_STRAIN_FLAG], 1);
}

constexpr int STRAIN_THREADS = 1024;
constexpr int STRAIN_TAB = STRAIN_THREADS;  // slots of a level on the LDS-mark path
constexpr int STRAIN_LDS_LEVELS = 65536;    // levels whose bits fit in LDS
constexpr int STRAIN_LW = STRAIN_LDS_LEVELS / 32;
constexpr int STRAIN_LP_MAX = 8192;         // levels whose slot offsets are kept in LDS

struct StrainCounts {
  int* flag;
  int* edges;
  int* degen;
};

__device__ inline bool bit_get(const uint32_t* b, int i) { return (b[i >> 5] >> (i & 31)) & 1u; }
__device__ inline void bit_set(uint32_t* b, int i) { atomicOr(&b[i >> 5], 1u << (i & 31)); }
// first set bit at index >= from (n if none)
__device__ inline int bit_next(const uint32_t* b, int from, int n) {
  int i = from;
  while (i < n) {
    const uint32_t w = b[i >> 5] >> (i & 31);
    if (w) return min(n, i + (int)__builtin_ctz(w));
    i = (i | 31) + 1;
  }
  return n;
}

// average the two velocities of an edge found over the limit
__device__ inline void strain_average(dvec4& a, dvec4& b, bool deg, int ia, int ib, dvec4* av,
                                      const StrainCounts& C) {
  double vt[3] = {a.x + b.x, a.y + b.y, a.z + b.z};
  vt[0] = 0.5 * vt[0];
  vt[1] = 0.5 * vt[1];
  vt[2] = 0.5 * vt[2];
  const dvec4 nv4 = dvec4{vt[0], vt[1], vt[2], 0.0};
  a = nv4;
  b = nv4;
  av[ia] = nv4;
  av[ib] = nv4;
  *C.flag = 1;
  atomicAdd(C.edges, 1);
  if (deg) atomicAdd(C.degen, 1);
}

// where a dirty vertex's next visit is marked
struct StrainMarks {
  uint32_t* tab_next;   // LDS marks of level L + 1 (fast path), indexed from nb
  int nb, ne;           // slot range of level L + 1; ne == nb when L + 1 is not on the fast path
  uint8_t* sflag_b;     // flag bytes in memory (other levels)
  const int4* snl;      // next-visit levels (for marks in memory)
  uint32_t* actb;       // LDS: level has work
  uint32_t* gscb;       // LDS: level must re-read its flag words
  bool use_lds;
};

// evaluate one slot's edges in order, (0,1), (1,2), (2,0), with the current
// velocities (reduceSuperelast's per-element loop, dcollid.cpp:515-560)
__device__ __attribute__((always_inline)) inline void strain_eval(const StrainHead& h, uint32_t f, const StrainBody* body,
                                   int s, int next_level, dvec4* av, const StrainMarks M,
                                   double dt, const StrainCounts C) {
  bool d0 = (f & 0xffu) != 0, d1 = (f & 0xff00u) != 0, d2 = (f & 0xff0000u) != 0;
  const bool tri = h.v[2] >= 0;
  // the whole body and all velocities are loaded before the first store
  // (the stores to av would otherwise order later body loads behind them)
  const StrainBody r = body[s];
  dvec4 a0 = av[h.v[0]], a1 = av[h.v[1]], a2{};
  if (tri) a2 = av[h.v[2]];
  bool g;
  if (edge_over_limit_l(r.x, r.x + 3, a0, a1, r.lold[0], r.L0[0], dt, &g)) {
    strain_average(a0, a1, g, h.v[0], h.v[1], av, C);
    d0 = d1 = true;
  }
  if (tri) {
    if (edge_over_limit_l(r.x + 3, r.x + 6, a1, a2, r.lold[1], r.L0[1], dt, &g)) {
      strain_average(a1, a2, g, h.v[1], h.v[2], av, C);
      d1 = d2 = true;
    }
    if (edge_over_limit_l(r.x + 6, r.x, a2, a0, r.lold[2], r.L0[2], dt, &g)) {
      strain_average(a2, a0, g, h.v[2], h.v[0], av, C);
      d2 = d0 = true;
    }
  }
  // every later visit of a dirty vertex must be evaluated
  const bool dk[3] = {d0, d1, d2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ns = h.ns[k];
    if (!dk[k] || ns < 0) continue;
    const int slot = ns >> 2;
    if (slot >= M.nb && slot < M.ne) {  // next visit on the following level
      atomicOr(&M.tab_next[slot - M.nb], 1u << (8 * (ns & 3)));
      if (M.use_lds) bit_set(M.actb, next_level);
    } else {
      M.sflag_b[ns] = 1;
      if (M.use_lds) {
        const int4 nl4 = M.snl[s];
        const int nl = k == 0 ? nl4.x : (k == 1 ? nl4.y : nl4.z);
        bit_set(M.actb, nl);
        bit_set(M.gscb, nl);
      }
    }
  }
}

// a head held in registers as two int4 (v0 v1 v2 ns0 | ns1 ns2 level elem),
// so a prefetched head never needs an addressable copy
struct HeadRegs {
  int4 a, b;
};
__device__ inline HeadRegs load_head(const StrainHead* __restrict__ head, int s) {
  const int4* p = reinterpret_cast<const int4*>(head + s);
  return HeadRegs{p[0], p[1]};
}
__device__ inline void store_head(StrainHead& h, const HeadRegs& r) {
  h.v[0] = r.a.x; h.v[1] = r.a.y; h.v[2] = r.a.z;
  h.ns[0] = r.a.w; h.ns[1] = r.b.x; h.ns[2] = r.b.y;
  h.level = r.b.z; h.elem = r.b.w;
}

// One work-group walks the levels in order.  On the fast path (levels of at
// most STRAIN_TAB slots) lane t owns slot b + t; larger levels are walked
// with a stride and always read their flag words from memory.
template <bool LP_LDS>
__global__ void __launch_bounds__(STRAIN_THREADS)
    k_strain_sweep(const int32_t* __restrict__ lvl_ptr, int n_levels,
                   const StrainHead* __restrict__ head, const int4* __restrict__ snl,
                   const StrainBody* __restrict__ body, uint32_t* sflag,
                   const uint8_t* __restrict__ lvl_active_g, dvec4* av, double dt,
                   int* __restrict__ counters) {
  __shared__ uint32_t actb[STRAIN_LW], preb[STRAIN_LW], gscb[STRAIN_LW];
  __shared__ uint32_t tab[2][STRAIN_TAB];
  __shared__ int32_t lp[STRAIN_LP_MAX + 1];
  // slots of the current level to evaluate, packed so that the FP64 work
  // of a level lands in as few wavefronts as possible
  __shared__ StrainHead wk_head[STRAIN_TAB];
  __shared__ int32_t wk_slot[STRAIN_TAB];
  __shared__ uint32_t wk_f[STRAIN_TAB];
  __shared__ int s_nwork[2];
  __shared__ int s_flag, s_edges, s_degen, s_state, s_levels, s_evals;
  const int t = threadIdx.x;
  if (t == 0) {
    s_state = counters[CNT_STRAIN_FLAG];  // <0 finished, 0 no edge over limit, >0 sweep needed
    s_flag = 0;
    s_edges = 0;
    s_degen = 0;
    s_levels = 0;
    s_evals = 0;
    s_nwork[0] = 0;
    s_nwork[1] = 0;
  }
  __syncthreads();
  const int state = s_state;
  if (state < 0) return;
  if (state == 0) {  // nothing over the limit: this sweep is a no-op and ends the loop
    if (t == 0) {
      counters[CNT_STRAIN_FLAG] = -1;
      counters[CNT_STRAIN_EDGES] = 0;
      atomicAdd(&counters[CNT_STRAIN_SWEEPS], 1);
    }
    return;
  }
  const StrainCounts C{&s_flag, &s_edges, &s_degen};
  const bool use_lds = n_levels <= STRAIN_LDS_LEVELS;
  const int nw = use_lds ? (n_levels + 31) / 32 : 0;
  for (int w = t; w < nw; w += STRAIN_THREADS) {
    uint32_t bits = 0;
    for (int j = 0; j < 32; ++j) {
      const int L = 32 * w + j;
      if (L < n_levels && lvl_active_g[L]) bits |= 1u << j;
    }
    actb[w] = bits;
    preb[w] = bits;
    gscb[w] = 0u;
  }
  for (int i = t; i < 2 * STRAIN_TAB; i += STRAIN_THREADS) (&tab[0][0])[i] = 0u;
  if (LP_LDS)
    for (int i = t; i <= n_levels; i += STRAIN_THREADS) lp[i] = lvl_ptr[i];
  __syncthreads();
  // slot offset of level i (from LDS when the offsets fit there)
  auto lvp = [&](int i) -> int {
    if constexpr (LP_LDS) return lp[i];
    else return lvl_ptr[i];
  };
  uint8_t* sflag_b = reinterpret_cast<uint8_t*>(sflag);
  // prefetched head + flag word of this lane's slot: level L + 1 (pn) and
  // the next pre-check-active level (pq)
  HeadRegs pn{}, pq{};
  uint32_t pnf = 0u, pqf = 0u;
  int pn_level = -1, pq_level = -1;
  int cur = 0, evals = 0;
  int L = use_lds ? bit_next(actb, 0, n_levels) : 0;
  while (L < n_levels) {
    const int b = lvp(L), e = lvp(L + 1);
    const bool fast = use_lds && (e - b) <= STRAIN_TAB;
    // slots of the following level (mark target of the fast path)
    int nb = 0, ne = 0, ne2 = 0;
    if (L + 1 < n_levels) {
      nb = e;
      ne2 = lvp(L + 2);
      ne = (use_lds && ne2 - nb <= STRAIN_TAB) ? ne2 : nb;
    }
    const StrainMarks M{tab[cur ^ 1], nb, ne, sflag_b, snl, actb, gscb, use_lds};
    if (fast) {
      HeadRegs h{};
      uint32_t f = 0u;
      const bool ok = b + t < e;
      if (pn_level == L) {
        h = pn;
        f = pnf;
      } else if (pq_level == L) {
        h = pq;
        f = pqf;
      } else if (ok) {
        h = load_head(head, b + t);
        f = sflag[b + t];
      }
      if (ok && bit_get(gscb, L)) f = sflag[b + t];  // marks written to memory since the prefetch
      const uint32_t m = tab[cur][t];
      f |= m;
      if (m) tab[cur][t] = 0u;
      if (ok && f != 0u) {
        const int k = atomicAdd(&s_nwork[cur], 1);
        store_head(wk_head[k], h);
        wk_slot[k] = b + t;
        wk_f[k] = f;
      }
      // prefetch the following level and the next pre-check-active level;
      // issued after this level's own loads were consumed, so no wait of
      // this level covers them
      pn_level = -1;
      if (ne > nb) {
        pn_level = L + 1;
        if (nb + t < ne) {
          pn = load_head(head, nb + t);
          pnf = sflag[nb + t];
        }
      }
      const int P = bit_next(preb, L + 2, n_levels);
      if (P < n_levels && P != pq_level) {
        const int pb = lvp(P), pe = lvp(P + 1);
        pq_level = -1;
        if (pe - pb <= STRAIN_TAB) {
          pq_level = P;
          if (pb + t < pe) {
            pq = load_head(head, pb + t);
            pqf = sflag[pb + t];
          }
        }
      }
      __syncthreads();
      const int nwork = s_nwork[cur];
      if (t == 0) s_nwork[cur ^ 1] = 0;  // next level's list; last read before the previous barrier
      for (int i = t; i < nwork; i += STRAIN_THREADS) {
        strain_eval(wk_head[i], wk_f[i], body, wk_slot[i], L + 1, av, M, dt, C);
        ++evals;
      }
    } else {
      // large level (or no LDS bits): flag words from memory, heads loaded here
      for (int s = b + t; s < e; s += STRAIN_THREADS) {
        const uint32_t f = sflag[s];
        if (f == 0u) continue;
        const StrainHead h = head[s];
        strain_eval(h, f, body, s, L + 1, av, M, dt, C);
        ++evals;
      }
      pn_level = -1;
    }
    if (t == 0) ++s_levels;
    __syncthreads();
    cur ^= 1;
    L = use_lds ? bit_next(actb, L + 1, n_levels) : L + 1;
  }
  if (evals) atomicAdd(&s_evals, evals);
  __syncthreads();
  if (t == 0) {
    counters[CNT_STRAIN_EDGES] = s_edges;
    counters[CNT_STRAIN_DEGEN] += s_degen;
    counters[CNT_STRAIN_SWEEPS] += 1;
    counters[CNT_STRAIN_LEVELS] += s_levels;
    counters[CNT_STRAIN_EVALS] += s_evals;
    counters[CNT_STRAIN_FLAG] = s_flag ? 0 : -1;  // next check decides; no edge -> finished
  }
}

hipError_t strain_limit(hipStream_t st, DevState& S, double dt, int max_sweeps, int* sweeps_out) {
  (void)sweeps_out;
  const int n = S.n_sched;
  if (S.ne == 0 || n == 0 || max_sweeps <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_strain_pack, dim3((n + 255) / 256), dim3(256), 0, st, S.shead, S.sbody,
                     S.x_old, n);
  for (int s = 0; s < max_sweeps; ++s) {
    CS_HIP_RET(hipMemsetAsync(S.lvl_active, 0, (size_t)S.n_levels, st));
    hipLaunchKernelGGL(k_strain_check, dim3((n + 255) / 256), dim3(256), 0, st, S.shead, S.sbody,
                       S.av, dt, n, S.sflag, S.lvl_active, S.counters);
    if (S.n_levels <= STRAIN_LP_MAX)
      hipLaunchKernelGGL(k_strain_sweep<true>, dim3(1), dim3(STRAIN_THREADS), 0, st, S.lvl_ptr,
                         S.n_levels, S.shead, S.snl, S.sbody, S.sflag, S.lvl_active, S.av, dt,
                         S.counters);
    else
      hipLaunchKernelGGL(k_strain_sweep<false>, dim3(1), dim3(STRAIN_THREADS), 0, st, S.lvl_ptr,
                         S.n_levels, S.shead, S.snl, S.sbody, S.sflag, S.lvl_active, S.av, dt,
                         S.counters);
  }
  return hipGetLastError();
}

}  // namespace cs
