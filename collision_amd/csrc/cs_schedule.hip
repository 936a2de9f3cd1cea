// cs_schedule.hip -- host-side level schedule of the strain-limit sweep.
//
// reduceSuperelastOnce (dcollid.cpp:486-560) visits the non-rigid elements in
// element order and averages the velocities of every edge over the strain
// limit in place.  An element's result depends on earlier elements only
// through shared vertices, so elements can be grouped into levels: an
// element's level is above the levels of all earlier elements that share a
// vertex with it.  Within a level no two elements share a vertex, and the
// elements of one vertex keep their order, so processing level by level
// gives the sequential sweep's results exactly.
//
// The device sweep (k_strain_pipe) runs several sweeps at once, sweep k+1
// trailing sweep k by `delta` levels.  That is exact when every vertex's
// visits lie within fewer than `delta` levels (its first visit in sweep k+1
// then comes after its last visit in sweep k).  The schedule is built so
// that this span is small:
//   1. earliest levels (1 + max level of earlier elements on shared vertices);
//   2. relaxation: a vertex's first visit is raised to at most W levels below
//      its last visit, repeated until nothing moves (W starts at the largest
//      vertex valence - 1 and doubles when a pass limit is hit; without
//      convergence the earliest levels are kept);
//   3. connected components much shorter than the schedule (strings, small
//      pieces) are shifted as a whole to the least loaded levels;
//   4. levels with more than STRAIN_LEVEL_CAP elements are split.
// Any schedule that keeps the element order of each vertex is exact; these
// steps only make the sweep faster.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "cs_internal.h"

namespace cs {

namespace {

int find_root(std::vector<int>& uf, int x) {
  while (uf[x] != x) {
    uf[x] = uf[uf[x]];
    x = uf[x];
  }
  return x;
}

}  // namespace

void build_strain_schedule(int nv, int ne, const int4* ev, const uint8_t* ef,
                           std::vector<int>& level, int& n_levels, int& delta) {
  level.assign((size_t)std::max(ne, 1), -1);
  n_levels = 0;
  delta = 1;
  auto np_of = [&](int e) { return (ef[e] & EF_TRI) ? 3 : 2; };
  auto vtx = [&](int e, int k) { return k == 0 ? ev[e].x : (k == 1 ? ev[e].y : ev[e].z); };
  int nsched = 0;
  std::vector<int> valence((size_t)nv, 0);
  for (int e = 0; e < ne; ++e) {
    if (ef[e] & EF_RIGID) continue;
    ++nsched;
    for (int k = 0; k < np_of(e); ++k) valence[vtx(e, k)]++;
  }
  if (nsched == 0) return;
  const int maxval = *std::max_element(valence.begin(), valence.end());

  // 1 + 2: earliest levels, relaxed so that each vertex's visits stay within W levels
  std::vector<int> lv((size_t)nv), last((size_t)nv), lev((size_t)ne, 0);
  std::vector<uint8_t> seen((size_t)nv);
  bool converged = false;
  for (int W = std::max(1, maxval - 1); W <= 64 && !converged; W *= 2) {
    std::fill(lev.begin(), lev.end(), 0);
    std::fill(last.begin(), last.end(), -1);
    for (int pass = 0; pass < 16; ++pass) {
      bool changed = false;
      std::fill(lv.begin(), lv.end(), -1);
      std::fill(seen.begin(), seen.end(), 0);
      for (int e = 0; e < ne; ++e) {
        if (ef[e] & EF_RIGID) continue;
        int l = lev[e];
        for (int k = 0; k < np_of(e); ++k) {
          const int v = vtx(e, k);
          l = std::max(l, lv[v] + 1);
          if (!seen[v]) l = std::max(l, last[v] - W);
        }
        if (l != lev[e]) {
          lev[e] = l;
          changed = true;
        }
        for (int k = 0; k < np_of(e); ++k) {
          lv[vtx(e, k)] = l;
          seen[vtx(e, k)] = 1;
        }
      }
      last = lv;
      if (!changed) {
        converged = true;
        break;
      }
    }
  }
  if (!converged) {  // earliest levels
    std::fill(lv.begin(), lv.end(), -1);
    for (int e = 0; e < ne; ++e) {
      if (ef[e] & EF_RIGID) continue;
      int l = 0;
      for (int k = 0; k < np_of(e); ++k) l = std::max(l, lv[vtx(e, k)] + 1);
      lev[e] = l;
      for (int k = 0; k < np_of(e); ++k) lv[vtx(e, k)] = l;
    }
  }
  int L = 0;
  for (int e = 0; e < ne; ++e)
    if (!(ef[e] & EF_RIGID)) L = std::max(L, lev[e] + 1);

  // 3: shift short components to the least loaded levels
  std::vector<int> uf((size_t)nv);
  std::iota(uf.begin(), uf.end(), 0);
  for (int e = 0; e < ne; ++e) {
    if (ef[e] & EF_RIGID) continue;
    int r0 = find_root(uf, vtx(e, 0));
    for (int k = 1; k < np_of(e); ++k) {
      const int r = find_root(uf, vtx(e, k));
      if (r == r0) continue;
      if (r < r0) {
        uf[r0] = r;
        r0 = r;
      } else {
        uf[r] = r0;
      }
    }
  }
  std::vector<int> root((size_t)ne, -1), cmin((size_t)nv, INT32_MAX), cmax((size_t)nv, -1);
  for (int e = 0; e < ne; ++e) {
    if (ef[e] & EF_RIGID) continue;
    const int r = find_root(uf, vtx(e, 0));
    root[e] = r;
    cmin[r] = std::min(cmin[r], lev[e]);
    cmax[r] = std::max(cmax[r], lev[e]);
  }
  auto is_short = [&](int r) { return 2 * (cmax[r] - cmin[r] + 1) <= L; };
  std::vector<int> load((size_t)L, 0);
  std::vector<int> shorts;
  std::vector<int> hist_off((size_t)nv, -1);
  for (int e = 0; e < ne; ++e) {
    if (root[e] < 0) continue;
    const int r = root[e];
    if (!is_short(r)) {
      load[lev[e]]++;
    } else if (hist_off[r] < 0) {
      hist_off[r] = 0;
      shorts.push_back(r);
    }
  }
  if (!shorts.empty()) {
    // per short component: element count of each of its levels
    size_t total = 0;
    for (int r : shorts) {
      hist_off[r] = (int)total;
      total += (size_t)(cmax[r] - cmin[r] + 1);
    }
    std::vector<int> hist(total, 0);
    for (int e = 0; e < ne; ++e)
      if (root[e] >= 0 && is_short(root[e])) hist[hist_off[root[e]] + lev[e] - cmin[root[e]]]++;
    std::vector<int> shift((size_t)nv, 0);
    for (int r : shorts) {
      const int len = cmax[r] - cmin[r] + 1, range = L - len;
      const int step = std::max(1, range / 128);
      const int* hr = hist.data() + hist_off[r];
      int best = 0, best_max = INT32_MAX;
      for (int o = 0; o <= range; o += step) {
        int m = 0;
        for (int j = 0; j < len && m < best_max; ++j) m = std::max(m, load[o + j] + hr[j]);
        if (m < best_max) {
          best_max = m;
          best = o;
        }
      }
      for (int j = 0; j < len; ++j) load[best + j] += hr[j];
      shift[r] = best - cmin[r];
    }
    for (int e = 0; e < ne; ++e)
      if (root[e] >= 0 && is_short(root[e])) lev[e] += shift[root[e]];
  }

  // 4: split levels above the cap (element order within a level)
  std::vector<int> cnt((size_t)L, 0);
  for (int e = 0; e < ne; ++e)
    if (root[e] >= 0) cnt[lev[e]]++;
  std::vector<int> base((size_t)L, 0), fill((size_t)L, 0);
  int nl = 0;
  for (int l = 0; l < L; ++l) {
    base[l] = nl;
    nl += std::max(1, (cnt[l] + STRAIN_LEVEL_CAP - 1) / STRAIN_LEVEL_CAP);
  }
  for (int e = 0; e < ne; ++e) {
    if (root[e] < 0) continue;
    const int l = lev[e];
    level[e] = base[l] + fill[l]++ / STRAIN_LEVEL_CAP;
  }
  n_levels = nl;

  // widest span of a vertex's visits
  std::vector<int> first((size_t)nv, -1);
  int span = 0;
  for (int e = 0; e < ne; ++e) {
    if (level[e] < 0) continue;
    for (int k = 0; k < np_of(e); ++k) {
      const int v = vtx(e, k);
      if (first[v] < 0) first[v] = level[e];
      span = std::max(span, level[e] - first[v]);
    }
  }
  delta = span + 1;
}

}  // namespace cs
