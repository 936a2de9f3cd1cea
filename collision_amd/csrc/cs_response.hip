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

// makeSet (dcollid.cpp:1015-1031) and createImpZoneForRG (dcollid3d.cpp:54-68):
// every point its own zone, every movable body one zone rooted at its
// smallest vertex
__global__ void k_uf_init(int32_t* __restrict__ uf, const int32_t* __restrict__ vbody,
                          const int32_t* __restrict__ bptr, const uint32_t* __restrict__ bmem,
                          int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int b = vbody ? vbody[v] : -1;
  uf[v] = b >= 0 ? (int32_t)bmem[bptr[b]] : v;
}

// Sort key of a hit: (lo, hi, sub-test) packed as tightly as the element
// count allows (hi in hb bits, the sub-test, at most 15, in 4 bits), which
// keeps the order of the record key and needs fewer digit passes.
__global__ void k_hit_keys(const hit_rec* __restrict__ hits, int n, int hb,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = hits[i].key;
  const uint64_t lo = k >> 36, hi = (k >> 8) & 0xFFFFFFFull, sub = k & 255u;
  keys[i] = (((lo << hb) | hi) << 4) | sub;
  idx[i] = (uint32_t)i;
}

// one record per (hit in canonical order, point): key = the vertex; the
// body spreads of an all-rigid contact use the slots of points 0 and 1 with
// key nv + body; points without a contribution get the sentinel nv + nbody
// and sort last
__global__ void k_expand_records(const hit_rec* __restrict__ hits,
                                 const uint32_t* __restrict__ sorted_idx, int n, int nv, int nbody,
                                 uint32_t* __restrict__ rkey, uint32_t* __restrict__ rval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * n) return;
  const int r = i >> 2, p = i & 3;
  const uint32_t h = sorted_idx[r];
  const int v = hits[h].v[p];
  const int b = p < 2 ? hits[h].sb[p] : -1;
  rkey[i] = (v >= 0) ? (uint32_t)v : (b >= 0 ? (uint32_t)(nv + b) : (uint32_t)(nv + nbody));
  rval[i] = h * 4u + (uint32_t)p;
}

// One thread per vertex segment of the sorted records.  The vertex joins the
// list of vertices whose average velocity this pass changed (vcount counts
// it; each vertex heads one segment).  A body segment (key nv + body) adds
// its spreads to the body's collsnImpulse_RG in canonical order, starting
// from the value it holds (the reference never clears it), and counts them.
__global__ void k_apply(const uint32_t* __restrict__ rkey, const uint32_t* __restrict__ rval,
                        int nrec, int nv, int nbody, const hit_rec* __restrict__ hits,
                        dvec4* __restrict__ av, uint8_t* __restrict__ has,
                        dvec4* __restrict__ rg_imp, int32_t* __restrict__ rg_num,
                        int32_t* __restrict__ vlist, int* __restrict__ vcount,
                        dvec4* __restrict__ pimp, int32_t* __restrict__ pnum,
                        int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = 0;
  bool head = false, bhead = false;
  if (i < nrec) {
    v = rkey[i];
    const bool first = !(i > 0 && rkey[i - 1] == v);
    head = first && v < (uint32_t)nv;  // head of the vertex segment
    bhead = first && v >= (uint32_t)nv && v < (uint32_t)(nv + nbody);
  }
  if (bhead) {
    const int b = (int)v - nv;
    dvec4 acc = rg_imp[b];
    int count = 0;
    for (int j = i; j < nrec && rkey[j] == v; ++j) {
      const uint32_t hv = rval[j];
      const hit_rec& h = hits[hv >> 2];
      const int p = hv & 3;
      acc.x = acc.x + h.imp[p][0];
      acc.y = acc.y + h.imp[p][1];
      acc.z = acc.z + h.imp[p][2];
      ++count;
    }
    rg_imp[b] = acc;
    rg_num[b] += count;
  }
  if (head) {
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
    if (pimp) {
      pimp[v] = dvec4{imp[0], imp[1], imp[2], 0.0};
      pnum[v] = count;
    }
    dvec4 a = av[v];
    a.x += (imp[0] + fric[0]) / count;
    a.y += (imp[1] + fric[1]) / count;
    a.z += (imp[2] + fric[2]) / count;
    if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) atomicOr(err, CS_DEVERR_NONFINITE);
    av[v] = a;
    has[v] = 1;
  }
  const int slot = wave_append(head, vcount);
  if (head) vlist[slot] = (int32_t)v;
}

static int bits_for(uint64_t maxval) {  // bits needed for values <= maxval, rounded to 8
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  b = (b + 7) & ~7;
  return b < 8 ? 8 : b;
}

hipError_t apply_hits(hipStream_t st, DevState& S, int nhits, int* vcount, bool record) {
  if (nhits <= 0) return hipSuccess;
  if (record) {  // a pass with hits replaces the recorded impulses
    CS_HIP_RET(hipMemsetAsync(S.pimp, 0, sizeof(dvec4) * (size_t)S.nv, st));
    CS_HIP_RET(hipMemsetAsync(S.pnum, 0, sizeof(int32_t) * (size_t)S.nv, st));
  }
  const int nb = (nhits + 255) / 256;
  int hb = 1;  // bits of an element id
  while (hb < 32 && ((uint64_t)S.ne >> hb) != 0) ++hb;
  hipLaunchKernelGGL(k_hit_keys, dim3(nb), dim3(256), 0, st, S.hits, nhits, hb, S.hkey, S.hidx);
  bool in_tmp = false;
  const int kbits = bits_for(((((uint64_t)S.ne << hb) | (uint64_t)S.ne) << 4) | 15u);
  CS_HIP_RET(radix_sort_u64(st, S.hkey, S.hidx, S.hkey_tmp, S.hidx_tmp, S.sort_hist, S.scan_part, nhits, kbits,
                            &in_tmp));
  const uint32_t* sidx = in_tmp ? S.hidx_tmp : S.hidx;
  const int nrec = 4 * nhits;
  hipLaunchKernelGGL(k_expand_records, dim3((nrec + 255) / 256), dim3(256), 0, st, S.hits, sidx,
                     nhits, S.nv, S.nbody, S.rkey, S.rval);
  CS_HIP_RET(radix_sort_u32(st, S.rkey, S.rval, S.rkey_tmp, S.rval_tmp, S.sort_hist, S.scan_part, nrec,
                            bits_for((uint64_t)S.nv + (uint64_t)S.nbody), &in_tmp));
  const uint32_t* rk = in_tmp ? S.rkey_tmp : S.rkey;
  const uint32_t* rv = in_tmp ? S.rval_tmp : S.rval;
  hipLaunchKernelGGL(k_apply, dim3((nrec + 255) / 256), dim3(256), 0, st, rk, rv, nrec, S.nv,
                     S.nbody, S.hits, S.av, S.has, S.rg_imp, S.rg_num, S.vlist, vcount,
                     record ? S.pimp : nullptr, S.pnum, S.counters + CNT_ERR);
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
// (also resets the zone count, which k_zone_heads fills after the sort)
__global__ void k_zone_keys(const uint32_t* __restrict__ zlist, int n, const int32_t* __restrict__ uf,
                            int vbits, uint64_t* __restrict__ key, uint32_t* __restrict__ val,
                            int* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) counters[CNT_ZONES] = 0;
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
constexpr int ZU = 2;  // member loads in flight per lane

__device__ inline void lane_total(double (*part)[9], int nout, double* tot) {
  for (int k = 0; k < nout; ++k) tot[k] = 0.0;
  for (int l = 0; l < ZONE_LANES; ++l)
    for (int k = 0; k < nout; ++k) tot[k] = tot[k] + part[l][k];
}

// Members: lane l takes members l, l + 64, ... of its zone in every phase.
// The first ZONE_CACHE members' positions, velocities, ids and fixed flags
// are kept in LDS by the lane that loaded them, so only the first phase
// reads them from memory; later members are read again.
constexpr int ZONE_CACHE = 512;

struct ZoneMember {
  dvec4 o, a;  // x_old, avgVel
  int p;       // vertex id
  bool fixed;
};

// Zones from the step's member list (zkey / zhead / zend, count in
// counters[CNT_ZONES]) or, with bptr, the movable bodies (members
// zval[bptr[z] .. bptr[z+1]), nbz zones).
__global__ void __launch_bounds__(ZONE_LANES)
    k_zone_update(const uint32_t* __restrict__ zkey, const uint32_t* __restrict__ zval,
                  const int32_t* __restrict__ zhead, const int32_t* __restrict__ zend,
                  const int32_t* __restrict__ bptr, int nbz,
                  const dvec4* __restrict__ x_old, dvec4* __restrict__ av,
                  const uint8_t* __restrict__ fixed, double m, double dt,
                  int* __restrict__ counters, const int32_t* __restrict__ vscene,
                  const int32_t* __restrict__ scene_active) {
  __shared__ double part[ZONE_LANES][9];
  __shared__ double cpos[ZONE_CACHE][3], cvel[ZONE_CACHE][3];
  __shared__ int cid[ZONE_CACHE];  // vertex id, -1 - id when fixed
  int* err = counters + CNT_ERR;
  const int l = threadIdx.x;
  const int nz = bptr ? nbz : counters[CNT_ZONES];
  for (int z = blockIdx.x; z < nz; z += gridDim.x) {
    const int b = bptr ? bptr[z] : zhead[z];
    const int e = bptr ? bptr[z + 1] : zend[zkey[b]];
    const int n = e - b;
    // batched scenes: only the zones of scenes still in their impact-zone
    // loop move (a repeated rigid update is not bitwise idempotent)
    if (vscene && !scene_active[vscene[zval[b]]]) continue;
    // member j (b <= j < e), from the cache when it holds it
    auto member = [&](int j) -> ZoneMember {
      ZoneMember r;
      const int c = j - b;
      if (c < ZONE_CACHE) {
        r.o = dvec4{cpos[c][0], cpos[c][1], cpos[c][2], 0.0};
        r.a = dvec4{cvel[c][0], cvel[c][1], cvel[c][2], 0.0};
        const int id = cid[c];
        r.fixed = id < 0;
        r.p = id < 0 ? -1 - id : id;
      } else {
        r.p = (int)zval[j];
        r.o = x_old[r.p];
        r.a = av[r.p];
        r.fixed = fixed[r.p] != 0;
      }
      return r;
    };
    // x_cm, v_cm (member loads ZU at a time, summed in member order)
    double t[9];
    for (int k = 0; k < 6; ++k) t[k] = 0.0;
    for (int j0 = b + l; j0 < e; j0 += ZU * ZONE_LANES) {
      dvec4 o[ZU], a[ZU];
      int pv[ZU];
      uint8_t fx[ZU];
#pragma unroll
      for (int u = 0; u < ZU; ++u) {
        const int j = j0 + u * ZONE_LANES;
        if (j < e) {
          const int p = zval[j];
          pv[u] = p;
          o[u] = x_old[p];
          a[u] = av[p];
          fx[u] = fixed[p];
        }
      }
#pragma unroll
      for (int u = 0; u < ZU; ++u) {
        const int j = j0 + u * ZONE_LANES;
        if (j >= e) break;
        t[0] = o[u].x + t[0]; t[1] = o[u].y + t[1]; t[2] = o[u].z + t[2];
        t[3] = a[u].x + t[3]; t[4] = a[u].y + t[4]; t[5] = a[u].z + t[5];
        const int c = j - b;
        if (c < ZONE_CACHE) {
          cpos[c][0] = o[u].x; cpos[c][1] = o[u].y; cpos[c][2] = o[u].z;
          cvel[c][0] = a[u].x; cvel[c][1] = a[u].y; cvel[c][2] = a[u].z;
          cid[c] = fx[u] ? -1 - pv[u] : pv[u];
        }
      }
    }
    for (int k = 0; k < 6; ++k) part[l][k] = t[k];
    __syncthreads();
    double tot[9];
    lane_total(part, 6, tot);
    const double x_cm[3] = {tot[0] / n, tot[1] / n, tot[2] / n};
    const double v_cm[3] = {tot[3] / n, tot[4] / n, tot[5] / n};
    __syncthreads();
    // angular momentum L = sum m dx x dv and inertia tensor
    // I = sum m (|dx|^2 delta - dx dx^T), in one sweep over the members
    double tl[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 9; ++k) t[k] = 0.0;
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const ZoneMember r = member(j);
      const double xo[3] = {r.o.x, r.o.y, r.o.z}, vv[3] = {r.a.x, r.a.y, r.a.z};
      double dx[3], dv[3], Li[3];
      sub3(xo, x_cm, dx);
      sub3(vv, v_cm, dv);
      cross3(dx, dv, Li);
      for (int k = 0; k < 3; ++k) tl[k] = m * Li[k] + tl[k];
      const double mag_dx = mag3(dx);
      for (int rr = 0; rr < 3; ++rr)
        for (int c = 0; c < 3; ++c) {
          double tm = -dx[rr] * dx[c];
          if (rr == c) tm += mag_dx * mag_dx;
          t[3 * rr + c] = tm * m + t[3 * rr + c];
        }
    }
    for (int k = 0; k < 3; ++k) part[l][k] = tl[k];
    __syncthreads();
    double L[3];
    lane_total(part, 3, L);
    __syncthreads();
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
    // the rotation of the zone, the same for every member
    const double cw = cs_det_cos(dt * mag_w);
    const double sw = mag_w < CS_ROUND_EPS ? 0.0 : cs_det_sin(dt * mag_w) / mag_w;
    for (int j = b + l; j < e; j += ZONE_LANES) {
      const ZoneMember r = member(j);
      if (r.fixed) continue;
      const double xo[3] = {r.o.x, r.o.y, r.o.z};
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
        for (int k = 0; k < 3; ++k) tmpV[k] = sw * w[k];
        cross3(tmpV, xR, wxR);
      }
      for (int k = 0; k < 3; ++k) {
        x_new[k] = x_cm[k] + dt * v_cm[k] + xF[k] + cw * xR[k] + wxR[k];
        an[k] = (x_new[k] - xo[k]) / dt;
      }
      if (isnan(an[0]) || isnan(an[1]) || isnan(an[2])) atomicOr(err, CS_DEVERR_NONFINITE);
      av[r.p] = dvec4{an[0], an[1], an[2], 0.0};
    }
    __syncthreads();  // the cache is refilled by the next zone
  }
}

// Every vertex of a zone with two or more members took part in a union, so
// the union calls list the members as they go (zone_touch); only the listed
// vertices are sorted, not the whole mesh.
hipError_t update_zones(hipStream_t st, DevState& S, double m, double dt, int nlist) {
  if (nlist <= 0) return hipMemsetAsync(S.counters + CNT_ZONES, 0, sizeof(int), st);  // no zones
  const int nb = (nlist + 255) / 256;
  int vbits = 1;
  while (vbits < 32 && ((uint64_t)(S.nv - 1) >> vbits) != 0) ++vbits;
  hipLaunchKernelGGL(k_zone_keys, dim3(nb), dim3(256), 0, st, S.zlist, nlist, S.uf, vbits,
                     S.zkey64, S.zval, S.counters);
  bool in_tmp = false;
  CS_HIP_RET(radix_sort_u64(st, S.zkey64, S.zval, S.zkey64_tmp, S.zval_tmp, S.sort_hist,
                            S.scan_part, nlist, bits_for(((uint64_t)(S.nv - 1) << vbits) | (uint64_t)(S.nv - 1)),
                            &in_tmp));
  const uint64_t* zk = in_tmp ? S.zkey64_tmp : S.zkey64;
  const uint32_t* zv = in_tmp ? S.zval_tmp : S.zval;
  hipLaunchKernelGGL(k_zone_heads, dim3(nb), dim3(256), 0, st, zk, nlist, vbits, S.zkey, S.zhead,
                     S.zend, S.counters);
  hipLaunchKernelGGL(k_zone_update, dim3(std::min(nlist, 2048)), dim3(ZONE_LANES), 0, st, S.zkey,
                     zv, S.zhead, S.zend, (const int32_t*)nullptr, 0, S.x_old, S.av, S.fixed, m, dt,
                     S.counters, S.nscene > 1 ? S.vscene : nullptr,
                     S.nscene > 1 ? S.scn + SCN_ACTIVE * S.nscene : nullptr);
  return hipGetLastError();
}

hipError_t update_body_zones(hipStream_t st, DevState& S, double m, double dt) {
  if (S.nbody <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zone_update, dim3(std::min(S.nbody, 2048)), dim3(ZONE_LANES), 0, st,
                     (const uint32_t*)nullptr, S.bmem, (const int32_t*)nullptr,
                     (const int32_t*)nullptr, S.bptr, S.nbody, S.x_old, S.av, S.fixed, m, dt,
                     S.counters, (const int32_t*)nullptr, (const int32_t*)nullptr);
  return hipGetLastError();
}

// The RG part of updateAverageVelocity (dcollid.cpp:728-734), one work-group
// per body: every point of a body with spreads this pass adds
// collsnImpulse_RG / collsn_num_RG (after its own impulses, applied by
// k_apply) and becomes has_collsn; the count is then cleared.
__global__ void k_body_rg_apply(const int32_t* __restrict__ bptr, const uint32_t* __restrict__ bmem,
                                dvec4* __restrict__ av, uint8_t* __restrict__ has,
                                const dvec4* __restrict__ rg_imp, int32_t* __restrict__ rg_num) {
  const int b = blockIdx.x;
  const int n = rg_num[b];
  if (n <= 0) return;  // uniform in the work-group
  const dvec4 g = rg_imp[b];
  for (int j = bptr[b] + threadIdx.x; j < bptr[b + 1]; j += blockDim.x) {
    const uint32_t v = bmem[j];
    dvec4 a = av[v];
    a.x += g.x / n;
    a.y += g.y / n;
    a.z += g.z / n;
    av[v] = a;
    has[v] = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) rg_num[b] = 0;
}

hipError_t apply_bodies(hipStream_t st, DevState& S, double m, double dt) {
  if (S.nbody <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_body_rg_apply, dim3(S.nbody), dim3(256), 0, st, S.bptr, S.bmem, S.av, S.has,
                     S.rg_imp, S.rg_num);
  if (dt > 0.0) return update_body_zones(st, S, m, dt);
  return hipGetLastError();
}

// updateFinalForRG (dcollid.cpp:626-675), one thread per body: the first
// point of the body (in element order) with has_collsn sets
// center_of_mass_velo = its avgVel and center_of_mass = avgVel dt + base,
// where base is the stored mrg_com when that point is the body's very first
// point and the current center_of_mass otherwise (an earlier point copied it
// into mrg_com); mrg_com then follows center_of_mass.
__global__ void k_final_bodies(const int32_t* __restrict__ bfirst, const int32_t* __restrict__ bvisit,
                               const dvec4* __restrict__ av, const uint8_t* __restrict__ has,
                               dvec4* __restrict__ bcom, dvec4* __restrict__ bvel,
                               dvec4* __restrict__ bmrg, double dt, int nbody) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbody) return;
  for (int j = bfirst[b]; j < bfirst[b + 1]; ++j) {
    const int v = bvisit[j];
    if (!has[v]) continue;
    const dvec4 base = (j == bfirst[b]) ? bmrg[b] : bcom[b];
    const dvec4 a = av[v];
    bvel[b] = dvec4{a.x, a.y, a.z, 0.0};
    bcom[b] = dvec4{a.x * dt + base.x, a.y * dt + base.y, a.z * dt + base.z, 0.0};
    break;
  }
  bmrg[b] = bcom[b];
}

void launch_final_bodies(hipStream_t st, DevState& S, double dt) {
  if (S.nbody <= 0) return;
  hipLaunchKernelGGL(k_final_bodies, dim3((S.nbody + 63) / 64), dim3(64), 0, st, S.bfirst, S.bvisit,
                     S.av, S.has, S.bcom, S.bvel, S.bmrg, dt, S.nbody);
}

// recordOriginPosition (dcollid.cpp:91-107)
__global__ void k_record_origin(const dvec4* __restrict__ x, dvec4* __restrict__ x_old,
                                uint8_t* __restrict__ has, const int32_t* __restrict__ vbody, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  has[v] = 0;
  if (vbody && vbody[v] >= 0) return;
  x_old[v] = x[v];
}

void launch_record_origin(hipStream_t st, DevState& S) {
  hipLaunchKernelGGL(k_record_origin, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.x, S.x_old,
                     S.has, S.vbody, S.nv);
}

__global__ void k_zone_labels(const int32_t* __restrict__ uf, int nv, int32_t* __restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) out[v] = uf_root(uf, v);
}

void launch_zone_labels(hipStream_t st, DevState& S, int32_t* out) {
  hipLaunchKernelGGL(k_zone_labels, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.uf, S.nv, out);
}

// ---------------------------------------------------------------------------
// detectDomainBoundaryCollision (dcollid.cpp:116-158): the reference visits a
// point once per element holding it (vadj_ptr gives that count).  A visit
// that leaves the average velocity unchanged leaves every later visit with
// the same input, so the loop stops there; a later visit can still change a
// point whose x_old lies outside the box, when the friction of an earlier
// visit shortens its inward move.  Movable rigid-body points are skipped.
__global__ void k_domain(dvec4* __restrict__ av, dvec4* __restrict__ x,
                         const dvec4* __restrict__ x_old, uint8_t* __restrict__ has,
                         const int32_t* __restrict__ vadj_ptr, const int32_t* __restrict__ vbody,
                         int nv, double lo0, double lo1, double lo2, double hi0, double hi1,
                         double hi2, double dt, double mu, double thickness, double mach_eps) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  if (vbody && vbody[v] >= 0) return;
  const int visits = vadj_ptr[v + 1] - vadj_ptr[v];
  dvec4 a4 = av[v], o4 = x_old[v], c4 = x[v];
  double a[3] = {a4.x, a4.y, a4.z}, o[3] = {o4.x, o4.y, o4.z}, c[3] = {c4.x, c4.y, c4.z};
  const double L[3] = {lo0, lo1, lo2}, U[3] = {hi0, hi1, hi2};
  bool touched = false;
  for (int visit = 0; visit < visits; ++visit) {
    const double a0 = a[0], a1 = a[1], a2 = a[2];
    double dv = 0;
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
    if (preVt > mach_eps) {
      const double f = smax(1.0 - mu * dv / preVt, 0.0);
#pragma unroll
      for (int j = 0; j < 3; ++j) a[j] *= f;
    }
    // bit patterns: -0.0 and 0.0 differ as results, so compare bits
    if (__double_as_longlong(a[0]) == __double_as_longlong(a0) &&
        __double_as_longlong(a[1]) == __double_as_longlong(a1) &&
        __double_as_longlong(a[2]) == __double_as_longlong(a2))
      break;
  }
  if (touched) has[v] = 1;
  av[v] = dvec4{a[0], a[1], a[2], 0.0};
  x[v] = dvec4{c[0], c[1], c[2], 0.0};
}

void launch_domain_boundary(hipStream_t st, DevState& S, const double lo[3], const double hi[3],
                            double dt, double mu, double thickness) {
  hipLaunchKernelGGL(k_domain, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.av, S.x, S.x_old,
                     S.has, S.vadj_ptr, S.vbody, S.nv, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], dt,
                     mu, thickness, CS_MACH_EPS);
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

// p[0..n) = v, 32-bit words (a kernel rather than hipMemsetAsync, so that a
// captured graph holds kernel nodes only)
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void launch_fill_u32(hipStream_t st, void* p, int64_t words, uint32_t v) {
  if (words <= 0) return;
  const int64_t blocks = std::min<int64_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, words, v);
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
  hipLaunchKernelGGL(k_uf_init, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.uf, S.vbody, S.bptr,
                     S.bmem, S.nv);
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

// Pipelined sparse exact sweeps.
//
// Sweeps run in one work-group, level by level; at iteration t sweep k works
// on level t - k * delta, where delta exceeds the level span of every
// vertex's visits.  A vertex's visits in sweep k + 1 therefore come after
// all its visits in sweep k, and every element sees the same velocities as
// in the sequential sweep loop.  Within one iteration the levels being
// processed share no vertex.
//
// A slot (element) is evaluated in sweep k only when its result can differ
// from "nothing over the limit":
//   * it wrote in sweep k - 1 (for the first sweep of a launch: the parallel
//     pre-check, on the velocities at the start of the sweep, found an edge
//     over the limit), or
//   * one of its vertices changed since its evaluation in sweep k - 1.
// Otherwise its inputs are the ones it saw, unchanged, in sweep k - 1, where
// nothing was over the limit.  Changes travel along each vertex's cycle of
// visits as a token: the visit that writes a vertex stores the new value
// into the next visit's velocity copy and marks it; a marked visit that does
// not write passes the value on.  The token ends where it started: the
// writer receives the value (refreshing its own copy) but is not marked, as
// it is evaluated anyway because it wrote.  A sweep in which nothing is over the
// limit ends the loop; it leaves no token and no written slot behind, so the
// sweeps after it evaluate nothing.
//
// Marks of the levels that may hold pending work (at most 64, as
// K * delta <= 64) live in an LDS ring, one byte per slot: bits 0-2 vertex
// changed, bit 3 wrote in the previous sweep.  With a single sweep in flight
// a mark 64 or more levels ahead goes to the slot's word in global memory.

// x_old of the slot vertices, edge lengths and velocity copies
__global__ void k_strain_pack(StrainSlot* __restrict__ slots, const dvec4* __restrict__ x_old,
                              const dvec4* __restrict__ av, int n, int with_x) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  StrainSlot& r = slots[s];
  const int v0 = r.v[0], v1 = r.v[1], v2 = r.v[2];
  const double none = __longlong_as_double(-1ll);
  if (with_x) {
    const dvec4 a = x_old[v0], b = x_old[v1];
    const double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
    r.x[0] = a.x; r.x[1] = a.y; r.x[2] = a.z;
    r.x[3] = b.x; r.x[4] = b.y; r.x[5] = b.z;
    r.lold[0] = dist3(pa, pb);
    if (v2 >= 0) {
      const dvec4 c = x_old[v2];
      const double pc[3] = {c.x, c.y, c.z};
      r.x[6] = c.x; r.x[7] = c.y; r.x[8] = c.z;
      r.lold[1] = dist3(pb, pc);
      r.lold[2] = dist3(pc, pa);
    }
  }
  const dvec4 a0 = av[v0], a1 = av[v1];
  r.vel[0] = dvec4{a0.x, a0.y, a0.z, none};
  r.vel[1] = dvec4{a1.x, a1.y, a1.z, none};
  if (v2 >= 0) {
    const dvec4 a2 = av[v2];
    r.vel[2] = dvec4{a2.x, a2.y, a2.z, none};
  }
}

// pre-check with the velocities at the start of the sweep: the slot's word
// (byte 3 set when an edge is over the limit, far marks cleared), lvl_active
// and the any-edge flag
__global__ void k_strain_check(const StrainSlot* __restrict__ slots, double dt, int n,
                               uint32_t* __restrict__ spre, uint8_t* __restrict__ lvl_active,
                               int* __restrict__ counters) {
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[CNT_STRAIN_PROG] = 0;
  if (counters[CNT_STRAIN_FLAG] < 0) return;  // loop already finished
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  bool any = false;
  if (s < n) {
    const StrainSlot& r = slots[s];
    const dvec4 a0 = r.vel[0], a1 = r.vel[1];
    bool deg;
    if (edge_over_limit_l(r.x, r.x + 3, a0, a1, r.lold[0], r.L0[0], dt, &deg)) any = true;
    if (r.v[2] >= 0) {
      const dvec4 a2 = r.vel[2];
      if (edge_over_limit_l(r.x + 3, r.x + 6, a1, a2, r.lold[1], r.L0[1], dt, &deg)) any = true;
      if (edge_over_limit_l(r.x + 6, r.x, a2, a0, r.lold[2], r.L0[2], dt, &deg)) any = true;
    }
    spre[s] = any ? 0x01000000u : 0u;
    if (any) lvl_active[r.level] = 1;
  }
  if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(&counters[CNT_STRAIN_FLAG], 1);
}

constexpr int SP_THREADS = 512;
constexpr int SP_RING = 64;            // levels of marks held in LDS
constexpr int SP_KMAX = 10;            // sweeps in flight
constexpr int SP_LP_MAX = 8192;        // levels whose slot offsets are kept in LDS
constexpr int SP_PRE_LEVELS = 65536;   // levels whose pre-check bits are kept in LDS
static_assert(STRAIN_LEVEL_CAP % 4 == 0 && STRAIN_LEVEL_CAP <= 1024, "work entry layout");

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

// velocity copy <-> (value, writer slot)
__device__ inline dvec4 vel_with_origin(const dvec4& a, int origin) {
  return dvec4{a.x, a.y, a.z, __longlong_as_double((long long)origin)};
}
__device__ inline int origin_of(const dvec4& a) { return (int)__double_as_longlong(a.w); }

// Every vertex's current velocity is in the copy of its last visit: a
// write's token reaches the last visit within the sweep of the write, and
// nothing changes the vertex after its last write.
__global__ void k_strain_writeback(const StrainSlot* __restrict__ slots,
                                   const int32_t* __restrict__ slast, dvec4* __restrict__ av,
                                   int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int lv = slast[v];
  if (lv < 0) return;
  const dvec4 c = slots[lv >> 2].vel[lv & 3];
  av[v] = dvec4{c.x, c.y, c.z, 0.0};
}

// sum over the wavefront
__device__ inline int wave_sum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// The pipe's records come from HBM: each level's slots were last touched by
// the pack and check kernels, which stream all of them.  A second work-group
// on the pipe's XCD (work-groups are dealt to the XCDs round-robin, so block
// kNumXcd shares block 0's) reads the records and pre-check words of the
// next SP_AHEAD levels into that XCD's L2 while the pipe works, following the
// iteration the pipe publishes.  It only reads, the pipe never waits for it,
// and it stops when the pipe is done, when it has read the last level, or
// after SP_IDLE polls without progress.
constexpr int SP_AHEAD = 8;
constexpr int SP_IDLE = 200000;

__device__ void strain_prefetch(const int32_t* __restrict__ lvl_ptr, int n_levels,
                                const StrainSlot* __restrict__ slots, const uint32_t* __restrict__ spre,
                                int* __restrict__ counters) {
  if (__hip_atomic_load(counters + CNT_STRAIN_FLAG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= 0)
    return;  // nothing to sweep (or the sweep already ended)
  const int tid = threadIdx.x;
  int next = 0, idle = 0;
  uint32_t acc = 0u;
  while (next < n_levels && idle < SP_IDLE) {
    const int p = __hip_atomic_load(counters + CNT_STRAIN_PROG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p < 0) break;
    if (p > next) next = min(p, n_levels);
    const int target = min(n_levels, p + SP_AHEAD);
    if (next >= target) {
      ++idle;
      __builtin_amdgcn_s_sleep(2);
      continue;
    }
    idle = 0;
    const int b = lvl_ptr[next], e = lvl_ptr[target];
    // one word of every 128 B line of the records, and of the pre-check words
    const uint32_t* rec = reinterpret_cast<const uint32_t*>(slots + b);
    for (int i = tid; i < 2 * (e - b); i += blockDim.x) acc += rec[32 * (size_t)i];
    for (int i = b + 32 * tid; i < e; i += 32 * (int)blockDim.x) acc += spre[i];
    next = target;
  }
  // (keeps the loads: the sum is never this value in practice, and a match
  // only writes an unused counter)
  if (acc == 0x9e3779b9u && tid == 0) counters[CNT_STRAIN_PROG + 1] = (int)acc;
}

// Work entry of one iteration: bits 0-9 slot index within its level, bits
// 10-13 sweep in flight, bits 14-17 marks.
template <bool LP_LDS>
__global__ void __launch_bounds__(SP_THREADS)
    k_strain_pipe(const int32_t* __restrict__ lvl_ptr, int n_levels, StrainSlot* slots,
                  uint32_t* spre, const uint8_t* __restrict__ lvl_active_g, double dt, int K,
                  int delta, int* __restrict__ counters) {
  static_assert(STRAIN_LEVEL_CAP % SP_THREADS == 0, "whole slots of a level per lane");
  constexpr int SPL = STRAIN_LEVEL_CAP / SP_THREADS;  // slots of a level per lane
  __shared__ uint32_t ring[SP_RING][STRAIN_LEVEL_CAP / 4];
  __shared__ uint32_t wk[SP_KMAX * STRAIN_LEVEL_CAP];
  __shared__ int32_t lp[LP_LDS ? SP_LP_MAX + 1 : 1];
  __shared__ uint32_t preb[SP_PRE_LEVELS / 32];
  __shared__ int s_nwork[2], s_last_mark, s_state;
  __shared__ uint32_t s_any[SP_THREADS / 64];
  __shared__ int s_last_edges[SP_THREADS / 64], s_degen[SP_THREADS / 64];
  if (blockIdx.x != 0) {  // the prefetcher (see strain_prefetch) and idle blocks
    if (blockIdx.x == kNumXcd) strain_prefetch(lvl_ptr, n_levels, slots, spre, counters);
    return;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    s_state = counters[CNT_STRAIN_FLAG];  // <0 finished, 0 no edge over limit, >0 sweeps needed
    s_nwork[0] = s_nwork[1] = 0;
    s_last_mark = INT32_MIN / 2;
  }
  __syncthreads();
  const int state = s_state;
  if (state < 0) return;
  if (state == 0) {  // nothing over the limit: this sweep is a no-op and ends the loop
    if (tid == 0) {
      counters[CNT_STRAIN_FLAG] = -1;
      counters[CNT_STRAIN_EDGES] = 0;
      atomicAdd(&counters[CNT_STRAIN_SWEEPS], 1);
    }
    return;
  }
  const bool use_preb = n_levels <= SP_PRE_LEVELS;
  for (int i = tid; i < SP_RING * STRAIN_LEVEL_CAP / 4; i += SP_THREADS) (&ring[0][0])[i] = 0u;
  if (use_preb)
    for (int w = tid; w < (n_levels + 31) / 32; w += SP_THREADS) {
      uint32_t bits = 0;
      for (int j = 0; j < 32; ++j) {
        const int L = 32 * w + j;
        if (L < n_levels && lvl_active_g[L]) bits |= 1u << j;
      }
      preb[w] = bits;
    }
  if constexpr (LP_LDS)
    for (int i = tid; i <= n_levels; i += SP_THREADS) lp[i] = lvl_ptr[i];
  __syncthreads();
  auto lvp = [&](int i) -> int {
    if constexpr (LP_LDS) return lp[i];
    else return lvl_ptr[i];
  };
  uint8_t* ring_b = reinterpret_cast<uint8_t*>(&ring[0][0]);
  uint8_t* spre_b = reinterpret_cast<uint8_t*>(spre);
  // The first sweep's pre-check results (and far marks) enter the ring one
  // level ahead: the word of this lane's slot in level l is loaded with the
  // slot records of an evaluation round, so waiting for it never waits for
  // stores (the vector memory counter retires loads and stores in order),
  // and the scan reads LDS only.
  struct PreWords {
    uint32_t w[SPL];
  };
  auto pre_word = [&](int l) -> PreWords {
    PreWords p;
#pragma unroll
    for (int j = 0; j < SPL; ++j) p.w[j] = 0u;
    if (l >= n_levels || (use_preb && !bit_get(preb, l))) return p;
    const int b = lvp(l), n = lvp(l + 1) - b;
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const int i = tid + j * SP_THREADS;
      if (i < n) p.w[j] = spre[b + i];
    }
    return p;
  };
  auto fold_pre = [&](int l, const PreWords& p) {
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      const uint32_t w = p.w[j];
      if (!w) continue;
      const int i = tid + j * SP_THREADS;
      const uint32_t f = ((w >> 24) ? 8u : 0u) | (w & 1u) | ((w >> 7) & 2u) | ((w >> 14) & 4u);
      atomicOr(&ring[l & (SP_RING - 1)][i >> 2], f << (8 * (i & 3)));
      if (w & 0x00ffffffu) spre[lvp(l) + i] = w & 0xff000000u;  // far marks consumed
    }
  };
  // per lane: sweeps with an edge over the limit (bit k), edges over the
  // limit in the last sweep in flight, degenerate edges
  uint32_t e_any = 0u;
  int e_last = 0, d_acc = 0;
  const int tend = n_levels + (K - 1) * delta;
  int t = use_preb ? bit_next(preb, 0, n_levels) : 0;
  int cur = 0, iters = 0;
  long long evals = 0;
  if (t < n_levels) fold_pre(t, pre_word(t));
  __syncthreads();
  while (t < tend) {
    if (tid == 0) __hip_atomic_store(counters + CNT_STRAIN_PROG, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // --- scan: lane tid looks at slots 2 tid, 2 tid + 1 of every level in
    // flight.  Level offsets: lane 2k (2k + 1) reads the first (last) offset
    // of sweep k's level, broadcast with readlane; then one 16-bit LDS read
    // per level.
    static_assert(SPL == 2, "two adjacent slots per lane");
    int lb[SP_KMAX], ln[SP_KMAX];
    {
      const int kk = lane >> 1, l = t - kk * delta;
      const bool ok = kk < K && l >= 0 && l < n_levels;
      const int v = ok ? lvp(l + (lane & 1)) : 0;
#pragma unroll
      for (int k = 0; k < SP_KMAX; ++k) {
        lb[k] = __builtin_amdgcn_readlane(v, 2 * k);
        ln[k] = __builtin_amdgcn_readlane(v, 2 * k + 1) - lb[k];
      }
    }
    uint32_t fk[SP_KMAX][SPL];
#pragma unroll
    for (int k = 0; k < SP_KMAX; ++k) {
      const int l = t - k * delta;
      const uint16_t* row =
          reinterpret_cast<const uint16_t*>(ring_b + (size_t)(l & (SP_RING - 1)) * STRAIN_LEVEL_CAP);
      const uint32_t w = 2 * tid < ln[k] ? row[tid] : 0u;
      fk[k][0] = w & 0xffu;
      fk[k][1] = 2 * tid + 1 < ln[k] ? (w >> 8) : 0u;
    }
    uint64_t mk[SP_KMAX][SPL];
    int total = 0;
#pragma unroll
    for (int k = 0; k < SP_KMAX; ++k) {
      const int l = t - k * delta;
      if (fk[k][0] | fk[k][1])
        reinterpret_cast<uint16_t*>(ring_b + (size_t)(l & (SP_RING - 1)) * STRAIN_LEVEL_CAP)[tid] = 0;
#pragma unroll
      for (int j = 0; j < SPL; ++j) {
        mk[k][j] = __ballot(fk[k][j] != 0u);
        total += __popcll(mk[k][j]);
      }
    }
    if (total) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&s_nwork[cur], total);
      base = __shfl(base, 0);
      const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
      for (int k = 0; k < SP_KMAX; ++k)
#pragma unroll
        for (int j = 0; j < SPL; ++j) {
          const int i = 2 * tid + j;
          if (fk[k][j])
            wk[base + __popcll(mk[k][j] & below)] = (uint32_t)i | ((uint32_t)k << 10) | (fk[k][j] << 14);
          base += __popcll(mk[k][j]);
        }
    }
    __syncthreads();
    // --- evaluate ---
    const int nwork = s_nwork[cur];
    if (tid == 0) s_nwork[cur ^ 1] = 0;  // next iteration's list
    bool marked = false;
    // one work entry: its slot record r, loaded by the caller
    auto evaluate = [&](const StrainSlot& r, uint32_t e) __attribute__((always_inline)) {
      const int idx = (int)(e & 1023u), k = (int)((e >> 10) & 15u);
      const uint32_t f = e >> 14;
      const int l = t - k * delta;
      const int s = lvp(l) + idx;
      const bool tri = r.v[2] >= 0;
      dvec4 a[3] = {r.vel[0], r.vel[1], tri ? r.vel[2] : dvec4{}};
      uint32_t wmask = 0u;
      int edges = 0, degen = 0;
      bool g;
      auto average = [&](int i, int j) {
        double vt[3] = {a[i].x + a[j].x, a[i].y + a[j].y, a[i].z + a[j].z};
        vt[0] = 0.5 * vt[0];
        vt[1] = 0.5 * vt[1];
        vt[2] = 0.5 * vt[2];
        a[i] = dvec4{vt[0], vt[1], vt[2], 0.0};
        a[j] = a[i];
        wmask |= (1u << i) | (1u << j);
        ++edges;
        if (g) ++degen;
      };
      if (edge_over_limit_l(r.x, r.x + 3, a[0], a[1], r.lold[0], r.L0[0], dt, &g)) average(0, 1);
      if (tri) {
        if (edge_over_limit_l(r.x + 3, r.x + 6, a[1], a[2], r.lold[1], r.L0[1], dt, &g))
          average(1, 2);
        if (edge_over_limit_l(r.x + 6, r.x, a[2], a[0], r.lold[2], r.L0[2], dt, &g))
          average(2, 0);
      }
      // every register of the record has arrived before the first store: a
      // load left outstanding (the third vertex of a bond) would make the
      // next iteration wait for this iteration's stores as well
      asm volatile("" ::"v"(r.x[6]), "v"(r.lold[2]), "v"(r.L0[2]), "v"(r.ns[2]), "v"(r.nl[2]),
                   "v"(r.vel[2].x), "v"(r.vel[2].w), "v"(r.vel[1].w), "v"(r.vel[0].w));
      const int np = tri ? 3 : 2;
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        if (p >= np) break;
        const bool wrote = (wmask >> p) & 1u;
        if (!wrote && !((f >> p) & 1u)) continue;
        const int ns = r.ns[p];
        if (ns < 0) continue;
        const int ts = ns >> 2, tp = ns & 3, tl = r.nl[p] >> 10, ti = r.nl[p] & 1023;
        const bool wrap = tl < l;  // the vertex's first visit, next sweep
        // no delivery comes back to this slot (a single visit, or the last
        // visit in the last sweep in flight): it keeps the value itself
        if (ts == s || (wrap && k == K - 1)) {
          if (wrote) slots[s].vel[p] = vel_with_origin(a[p], s);
          continue;
        }
        const int origin = wrote ? s : origin_of(r.vel[p]);
        slots[ts].vel[tp] = vel_with_origin(a[p], origin);
        // back at the writer: delivered, not marked (it is evaluated anyway,
        // because it wrote)
        if (ts == origin) continue;
        if (!wrap && tl - l >= SP_RING) {       // far ahead (one sweep in flight)
          spre_b[4 * (size_t)ts + tp] = 1;
          if (use_preb) bit_set(preb, tl);
        } else {
          atomicOr(&ring[tl & (SP_RING - 1)][ti >> 2], (1u << tp) << (8 * (ti & 3)));
          marked = true;
        }
      }
      if (edges) {
        e_any |= 1u << k;
        if (k == K - 1) e_last += edges;
        d_acc += degen;
        if (k < K - 1) {  // evaluate again in the next sweep
          atomicOr(&ring[l & (SP_RING - 1)][idx >> 2], 8u << (8 * (idx & 3)));
          marked = true;
        }
      }
    };
    auto slot_of_entry = [&](uint32_t e) {
      return lvp(t - (int)((e >> 10) & 15u) * delta) + (int)(e & 1023u);
    };
    // first round: this lane's entry and the first sweep's next pre-check
    // word are loaded together, the word folded into the ring before any
    // store of this iteration
    const PreWords pw = pre_word(t + 1);
    const bool have = tid < nwork;
    const uint32_t e0 = have ? wk[tid] : 0u;
    StrainSlot r0;
    if (have) r0 = slots[slot_of_entry(e0)];
    fold_pre(t + 1, pw);  // (no word beyond the last level)
    if (have) evaluate(r0, e0);
    for (int q = tid + SP_THREADS; q < nwork; q += SP_THREADS) {
      const uint32_t e = wk[q];
      const StrainSlot r = slots[slot_of_entry(e)];
      evaluate(r, e);
    }
    // every mark is consumed within delta iterations
    if (__ballot(marked) && lane == 0) s_last_mark = t;
    if (tid == 0) evals += nwork;
    ++iters;
    __syncthreads();
    cur ^= 1;
    int tn = t + 1;
    if (use_preb && t >= s_last_mark + delta) {
      // nothing pending: only the first sweep's pre-check levels remain
      tn = bit_next(preb, t + 1, n_levels);
      if (tn >= n_levels) break;
    }
    if (tn != t + 1) {  // skipped: the pre-check word of the new level
      fold_pre(tn, pre_word(tn));
      __syncthreads();
    }
    t = tn;
  }
  if (tid == 0) __hip_atomic_store(counters + CNT_STRAIN_PROG, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // totals
  {
    uint32_t any = e_any;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) any |= __shfl_xor(any, o);
    const int el = wave_sum(e_last), dg = wave_sum(d_acc);
    if (lane == 0) {
      s_any[wave] = any;
      s_last_edges[wave] = el;
      s_degen[wave] = dg;
    }
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t any = 0u;
    int last = 0, degen = 0;
    for (int w = 0; w < SP_THREADS / 64; ++w) {
      any |= s_any[w];
      last += s_last_edges[w];
      degen += s_degen[w];
    }
    int nz = 0;
    while (nz < K && ((any >> nz) & 1u)) ++nz;
    const int niter = nz < K ? nz + 1 : K;  // sweeps the sequential loop runs
    counters[CNT_STRAIN_EDGES] = nz < K ? 0 : last;
    counters[CNT_STRAIN_DEGEN] += degen;
    counters[CNT_STRAIN_SWEEPS] += niter;
    counters[CNT_STRAIN_LEVELS] += iters;
    counters[CNT_STRAIN_EVALS] += (int)evals;
    counters[CNT_STRAIN_FLAG] = (nz == K) ? 0 : -1;  // all K wrote: the loop goes on
  }
}

hipError_t strain_limit(hipStream_t st, DevState& S, double dt, int max_sweeps,
                        std::vector<hipEvent_t>* marks) {
  const int n = S.n_sched;
  if (S.ne == 0 || n == 0 || max_sweeps <= 0) return hipSuccess;
  // sweeps in flight: marks pending at one time cover K * delta - 1 levels
  const int kfit = std::max(1, SP_RING / std::max(1, S.strain_delta));
  for (int k0 = 0; k0 < max_sweeps;) {
    const int K = std::min({SP_KMAX, kfit, max_sweeps - k0});
    hipLaunchKernelGGL(k_strain_pack, dim3((n + 255) / 256), dim3(256), 0, st, S.sslot, S.x_old,
                       S.av, n, k0 == 0 ? 1 : 0);
    CS_HIP_RET(hipMemsetAsync(S.lvl_active, 0, (size_t)S.n_levels, st));
    hipLaunchKernelGGL(k_strain_check, dim3((n + 255) / 256), dim3(256), 0, st, S.sslot, dt, n,
                       S.spre, S.lvl_active, S.counters);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (marks && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess)
      (void)hipEventRecord(e0, st);
    if (S.n_levels <= SP_LP_MAX)
      hipLaunchKernelGGL(k_strain_pipe<true>, dim3(kNumXcd + 1), dim3(SP_THREADS), 0, st, S.lvl_ptr,
                         S.n_levels, S.sslot, S.spre, S.lvl_active, dt, K, S.strain_delta,
                         S.counters);
    else
      hipLaunchKernelGGL(k_strain_pipe<false>, dim3(kNumXcd + 1), dim3(SP_THREADS), 0, st, S.lvl_ptr,
                         S.n_levels, S.sslot, S.spre, S.lvl_active, dt, K, S.strain_delta,
                         S.counters);
    if (e1) {
      (void)hipEventRecord(e1, st);
      marks->push_back(e0);
      marks->push_back(e1);
    }
    hipLaunchKernelGGL(k_strain_writeback, dim3((S.nv + 255) / 256), dim3(256), 0, st, S.sslot,
                       S.slast, S.av, S.nv);
    k0 += K;
  }
  return hipGetLastError();
}

}  // namespace cs
