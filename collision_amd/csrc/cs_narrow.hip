// Narrow phase for gfx950: static proximity and continuous collision tests
// over the compacted candidate stream, FP64.
//
// One lane per (pair, sub-test): each lane gathers its sub-test's four
// points (x_old, avgVel, static flag) and runs it.  Sub-tests follow the
// reference exactly:
//   proximity: TriToTri / TriToBond / BondToBond (dcollid3d.cpp:485-627),
//              PointToTri (778-922), EdgeToEdge (643-776) with h = thickness
//   CCD:       MovingTriToTri / MovingTriToBond / MovingBondToBond
//              (dcollid3d.cpp:203-325) -> MovingPointToTri / MovingEdgeToEdge
//              (327-369) -> isCoplanar (371-482), h = rounding tolerance
//   response:  PointToTriImpulse (925-1107), EdgeToEdgeImpulse (1109-1300)
// Instead of accumulating into per-vertex state, every hit writes a record
// with its per-point impulse / friction increments; the response stage sums
// them per vertex in canonical order, reproducing the reference's
// sequential accumulation bit for bit.
// In impact-zone passes the points of every sub-test from a pair's first
// hit onward are merged (the reference's cumulative-status rule,
// dcollid3d.cpp:222-229, 300-302, 319-321) by a lock-free union-find whose
// roots are the smallest vertex id of each zone.
#include <algorithm>

#include "cs_internal.h"
#include "cs_mlp.h"

namespace cs {

constexpr int SUB_GROUP = 16;  // lanes per pair (sub-tests 0..14)
// learned form: codes of the per-slot seeds (a scored slot holds its seed in
// [0, 1], sign bit set when a root is predicted) and the 8-bit pair keys
constexpr float SEED_CERTAIN = 4.0f, SEED_INACTIVE = 8.0f;
constexpr uint32_t KEY_NONE = 255u;  // 0 = best score .. 254; 255 = no scored slot

struct Sub {
  double xo[4][3];
  double v[4][3];
  bool st[4];   // static (is_fixed)
  bool mv[4];   // movable rigid body (is_movableRG)
  int bd[4];    // body of a movable point
};

// Per-point impulse and friction increments of one hit.  A contact whose
// four points are all rigid adds nothing to its points (all_rigid) and
// spreads imp[0] over body sb[0], imp[1] over body sb[1] instead.
struct Contrib {
  double imp[4][3];
  double fric[4][3];
  bool all_rigid;
  int sb[2];
};

__device__ inline bool rigid_pt(const Sub& S, int q) { return S.st[q] || S.mv[q]; }

// all-rigid contact: SpreadImpactZoneImpulse(pts[a], rigid_impulse[0]) and
// (pts[b], -rigid_impulse[1]) for the movable ones (dcollid3d.cpp:1043-1052,
// 1225-1234)
__device__ inline void spread(const Sub& S, int a, int b, double r0, double r1, const double nor[3],
                              Contrib& C) {
  C.all_rigid = true;
  C.sb[0] = S.mv[a] ? S.bd[a] : -1;
  C.sb[1] = S.mv[b] ? S.bd[b] : -1;
  const double i1 = -1.0 * r1;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    C.imp[0][j] = r0 * nor[j];
    C.imp[1][j] = i1 * nor[j];
#pragma unroll
    for (int q = 2; q < 4; ++q) C.imp[q][j] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) C.fric[q][j] = 0.0;
  }
}

// PointToTriImpulse (dcollid3d.cpp:925-1107)
__device__ void d_pt_impulse(const Sub& S, const NarrowParams& P, const double nor[3], double w[3],
                             double dist, Contrib& C, int* err) {
  double v_rel[3] = {0.0, 0.0, 0.0}, vn, vt;
  double impulse = 0.0, m_impulse = 0.0, sum_w = 0.0;
  double rig[2] = {0.0, 0.0};
  const double k = P.k, m = P.m, dt = P.dt, lambda = P.lambda, h = P.thickness;
  dist = h - dist;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v_rel[i] += S.v[3][i];
#pragma unroll
    for (int j = 0; j < 3; ++j) v_rel[i] -= w[j] * S.v[j][i];
  }
  vn = dot3(v_rel, nor);
  if (dot3(v_rel, v_rel) > vn * vn) vt = sqrt(dot3(v_rel, v_rel) - vn * vn);
  else vt = 0.0;
  if (vn < 0) {
    if (S.st[3] || (S.st[0] && S.st[1] && S.st[2])) {
      impulse = vn;
      rig[0] = vn;
      rig[1] = vn;
    } else if (S.mv[0] && S.mv[1] && S.mv[2] && S.mv[3]) {
      const double m1 = P.bmass[S.bd[0]], m2 = P.bmass[S.bd[3]];
      rig[0] = vn * m2 / (m1 + m2);
      rig[1] = vn * m1 / (m1 + m2);
    } else if (S.mv[0] && S.mv[1] && S.mv[2]) {
      rig[0] = 0.5 * vn;
      impulse = 0.5 * vn;
    } else if (S.mv[3]) {
      impulse = 0.5 * vn;
      rig[1] = 0.5 * vn;
    } else {
      impulse = vn * 0.5;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (S.st[i]) w[i] = 0.0;
      sum_w += w[i];
    }
    if (fabs(sum_w) > CS_MACH_EPS) {
      const double inv = 1.0 / sum_w;
#pragma unroll
      for (int i = 0; i < 3; ++i) w[i] = inv * w[i];
    }
  }
  const bool all_rigid = rigid_pt(S, 0) && rigid_pt(S, 1) && rigid_pt(S, 2) && rigid_pt(S, 3);
  if (vn * dt < 0.1 * dist) {
    if (all_rigid) {
      rig[0] *= 1.0 + P.cr;
      rig[1] *= 1.0 + P.cr;
    } else {
      const double tmp = -smin(dt * k * dist / m, (0.1 * dist / dt - vn));
      impulse += tmp;
      rig[0] += tmp;
      rig[1] += tmp;
    }
  }
  if (fabs(sum_w) < CS_MACH_EPS) m_impulse = impulse;
  else m_impulse = 2.0 * impulse / (1.0 + dot3(w, w));
  if (all_rigid) {
    spread(S, 0, 3, rig[0], rig[1], nor, C);
    return;
  }
  C.all_rigid = false;
  C.sb[0] = C.sb[1] = -1;
  const bool fr = fabs(vt) > CS_ROUND_EPS;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double t_impulse = S.mv[i] ? rig[0] : m_impulse;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      C.imp[i][j] = w[i] * t_impulse * nor[j];
      C.fric[i][j] = fr ? smax(-fabs(lambda * w[i] * t_impulse / vt), -1.0) * (v_rel[j] - vn * nor[j]) : 0.0;
    }
  }
  {
    const double t_impulse = S.mv[3] ? rig[1] : m_impulse;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      C.imp[3][j] = -(t_impulse * nor[j]);
      C.fric[3][j] = fr ? smax(-fabs(lambda * t_impulse / vt), -1.0) * (v_rel[j] - vn * nor[j]) : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (!S.st[q] && !isfinite(C.imp[q][j])) atomicOr(err, CS_DEVERR_NONFINITE);
}

// EdgeToEdgeImpulse (dcollid3d.cpp:1109-1300)
__device__ void d_ee_impulse(const Sub& S, const NarrowParams& P, const double nor[3], double a,
                             double b, double dist, Contrib& C, int* err) {
  double v_rel[3], vn, vt, impulse = 0.0, m_impulse = 0.0;
  double rig[2] = {0.0, 0.0};
  const double k = P.k, m = P.m, dt = P.dt, lambda = P.lambda, h = P.thickness;
  dist = h - dist;
  double wa0 = 1.0 - a, wa1 = a, wb0 = 1.0 - b, wb1 = b;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    v_rel[j] = (1.0 - b) * S.v[2][j] + b * S.v[3][j];
    v_rel[j] -= (1.0 - a) * S.v[0][j] + a * S.v[1][j];
  }
  vn = dot3(v_rel, nor);
  if (dot3(v_rel, v_rel) > vn * vn) vt = sqrt(dot3(v_rel, v_rel) - vn * vn);
  else vt = 0.0;
  if (vn < 0.0) {
    if ((S.st[0] && S.st[1]) || (S.st[2] && S.st[3])) {
      impulse = vn;
      rig[0] = vn;
      rig[1] = vn;
    } else if (S.mv[0] && S.mv[1] && S.mv[2] && S.mv[3]) {
      const double m1 = P.bmass[S.bd[0]], m2 = P.bmass[S.bd[2]];
      rig[0] = vn * m2 / (m1 + m2);
      rig[1] = vn * m1 / (m1 + m2);
    } else if (S.mv[0] && S.mv[1]) {
      rig[0] = 0.5 * vn;
      impulse = 0.5 * vn;
    } else if (S.mv[2] && S.mv[3]) {
      impulse = 0.5 * vn;
      rig[1] = 0.5 * vn;
    } else {
      impulse = vn * 0.5;
    }
    if (S.st[0]) wa0 = 0.0;
    if (S.st[1]) wa1 = 0.0;
    if (S.st[2]) wb0 = 0.0;
    if (S.st[3]) wb1 = 0.0;
  }
  const bool all_rigid = rigid_pt(S, 0) && rigid_pt(S, 1) && rigid_pt(S, 2) && rigid_pt(S, 3);
  if (vn * dt < 0.1 * dist) {
    if (all_rigid) {
      rig[0] *= 1.0 + P.cr;
      rig[1] *= 1.0 + P.cr;
    } else {
      const double tmp = -smin(dt * k * dist / m, (0.1 * dist / dt - vn));
      impulse += tmp;
      rig[0] += tmp;
      rig[1] += tmp;
    }
  }
  if (wa0 + wa1 < CS_MACH_EPS || wb0 + wb1 < CS_MACH_EPS) m_impulse = impulse;
  else m_impulse = 2.0 * impulse / (wa0 * wa0 + wa1 * wa1 + wb0 * wb0 + wb1 * wb1);
  if (all_rigid) {
    spread(S, 0, 2, rig[0], rig[1], nor, C);
    return;
  }
  C.all_rigid = false;
  C.sb[0] = C.sb[1] = -1;
  const bool fr = fabs(vt) > CS_ROUND_EPS;
  const double wgt[4] = {wa0, wa1, wb0, wb1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double t_impulse = S.mv[q] ? rig[q < 2 ? 0 : 1] : m_impulse;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double d = wgt[q] * t_impulse * nor[j];
      C.imp[q][j] = (q < 2) ? d : -d;
      C.fric[q][j] = fr ? smax(-fabs(lambda * wgt[q] * t_impulse / vt), -1.0) * (v_rel[j] - vn * nor[j]) : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (!S.st[q] && !isfinite(C.imp[q][j])) atomicOr(err, CS_DEVERR_NONFINITE);
}

// PointToTri (dcollid3d.cpp:778-922): c = coords (triangle 0..2, point 3)
__device__ bool d_point_to_tri(const Sub& S, const double c[4][3], double h, const NarrowParams& P,
                               Contrib& C, int* err) {
  double w[3] = {0.0, 0.0, 0.0};
  double x13[3], x23[3], x43[3], nor[3], nor_mag, dist, det;
  sub3(c[0], c[2], x13);
  sub3(c[1], c[2], x23);
  sub3(c[3], c[2], x43);
  det = dot3(x13, x13) * dot3(x23, x23) - dot3(x13, x23) * dot3(x13, x23);
  if (fabs(det) < 1000 * CS_MACH_EPS) return false;
  cross3(x13, x23, nor);
  nor_mag = mag3(nor);
  double x43_old[3];
  sub3(S.xo[3], S.xo[2], x43_old);
  dist = dot3(x43_old, nor);
  const double sgn = (dist >= 0) ? 1.0 : -1.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) nor[i] /= nor_mag * sgn;
  dist = fabs(dot3(x43, nor));
  w[0] = (dot3(x13, x43) * dot3(x23, x23) - dot3(x23, x43) * dot3(x13, x23)) / det;
  w[1] = (dot3(x13, x13) * dot3(x23, x43) - dot3(x13, x23) * dot3(x13, x43)) / det;
  w[2] = 1 - w[0] - w[1];
  if (fabs(w[0]) < CS_ROUND_EPS || fabs(w[1]) < CS_ROUND_EPS || fabs(w[2]) < CS_ROUND_EPS) {
    double vec[3] = {S.xo[3][0], S.xo[3][1], S.xo[3][2]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) vec[j] -= w[i] * S.xo[i][j];
    if (mag3(vec) > CS_ROUND_EPS) {
      nor[0] = vec[0];
      nor[1] = vec[1];
      nor[2] = vec[2];
    }
  }
  nor_mag = mag3(nor);
  if (nor_mag > CS_ROUND_EPS) {
#pragma unroll
    for (int i = 0; i < 3; ++i) nor[i] /= nor_mag;
  } else {
    atomicOr(err, CS_DEVERR_NONFINITE);
    return false;
  }
  if (dist > h) return false;
  const double eps = P.round_eps;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (w[i] > 1 + eps || w[i] < -eps) return false;
  d_pt_impulse(S, P, nor, w, dist, C, err);
  return true;
}

// EdgeToEdge (dcollid3d.cpp:643-776): edges c0c1 and c2c3
__device__ bool d_edge_to_edge(const Sub& S, const double c[4][3], double h, const NarrowParams& P,
                               Contrib& C, int* err) {
  double x21[3], x43[3], x31[3], tmp[3], v1[3], v2[3], nor[3], nor_mag, dist, a, b;
  sub3(c[1], c[0], x21);
  sub3(c[3], c[2], x43);
  sub3(c[2], c[0], x31);
  cross3(x21, x43, tmp);
  if (mag3(tmp) < CS_ROUND_EPS) return false;
  a = (dot3(x43, x43) * dot3(x21, x31) - dot3(x21, x43) * dot3(x43, x31)) /
      (dot3(x21, x21) * dot3(x43, x43) - dot3(x21, x43) * dot3(x21, x43));
  b = (dot3(x21, x43) * dot3(x21, x31) - dot3(x21, x21) * dot3(x43, x31)) /
      (dot3(x21, x21) * dot3(x43, x43) - dot3(x21, x43) * dot3(x21, x43));
  a = smax(smin(a, 1.0), 0.0);
  b = smax(smin(b, 1.0), 0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = a * x21[i];
    v2[i] = b * x43[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = c[0][i] + v1[i];
    v2[i] = c[2][i] + v2[i];
  }
  sub3(v2, v1, nor);
  nor_mag = mag3(nor);
  if (nor_mag < 1000 * CS_MACH_EPS) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      nor[j] = (1.0 - b) * S.xo[2][j] + b * S.xo[3][j];
      nor[j] -= (1.0 - a) * S.xo[0][j] + a * S.xo[1][j];
    }
  }
  dist = dist3(v1, v2);
  if (dist > h) return false;
  nor_mag = mag3(nor);
  if (nor_mag < CS_MACH_EPS) {
    atomicOr(err, CS_DEVERR_NONFINITE);
    return false;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) nor[i] /= nor_mag;
  d_ee_impulse(S, P, nor, a, b, dist, C, err);
  return true;
}

// Certified screen of isCoplanar: true when isCoplanar (below) certainly
// returns false.  The cubic's coefficients, Q, R and every branch decision are
// computed exactly as below; the roots are then taken with the fast
// acos / cos / cube root (a few ulps; the cube root also stands for
// pow(x, 1.0/3.0), which differs from it by at most 1.85e-17 |ln x|
// relative).  Each approximate root is within margin = 1e-13 (|scale| + dt)
// of the correctly rounded one, about a hundred times the worst error of
// those approximations; a root counts as certainly invalid only when it lies
// beyond the validity window (2 MACH_EPS, dt + MACH_EPS] by more than the
// margin, and the Cardano branch's two thresholds on |A| and |A - B| must be
// cleared by the margin as well.  Anything else -- NaN included -- is not
// certain and runs the exact path.  When |a| <= MACH_EPS (the quadratic and
// linear branch, no transcendental) the screen evaluates isCoplanar exactly.
__device__ bool d_coplanar_certainly_false(const Sub& S, double dt) {
  double v[4][3], x[4][3];
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[i][k] = S.v[i][k] - S.v[0][k];
      x[i][k] = S.xo[i][k] - S.xo[0][k];
    }
  double a, b, c, d, vv[3], vx[3], xx[3];
  vv[0] = v[1][1] * v[2][2] - v[1][2] * v[2][1];
  vv[1] = v[1][0] * v[2][2] - v[1][2] * v[2][0];
  vv[2] = v[1][0] * v[2][1] - v[1][1] * v[2][0];
  vx[0] = v[1][1] * x[2][2] - v[1][2] * x[2][1] - v[2][1] * x[1][2] + v[2][2] * x[1][1];
  vx[1] = v[1][0] * x[2][2] - v[1][2] * x[2][0] - v[2][0] * x[1][2] + v[2][2] * x[1][0];
  vx[2] = v[1][0] * x[2][1] - v[1][1] * x[2][0] - v[2][0] * x[1][1] + v[2][1] * x[1][0];
  xx[0] = x[1][1] * x[2][2] - x[1][2] * x[2][1];
  xx[1] = x[1][0] * x[2][2] - x[1][2] * x[2][0];
  xx[2] = x[1][0] * x[2][1] - x[1][1] * x[2][0];
  a = v[3][0] * vv[0] - v[3][1] * vv[1] + v[3][2] * vv[2];
  b = x[3][0] * vv[0] - x[3][1] * vv[1] + x[3][2] * vv[2] + v[3][0] * vx[0] - v[3][1] * vx[1] +
      v[3][2] * vx[2];
  c = x[3][0] * vx[0] - x[3][1] * vx[1] + x[3][2] * vx[2] + v[3][0] * xx[0] - v[3][1] * xx[1] +
      v[3][2] * xx[2];
  d = x[3][0] * xx[0] - x[3][1] * xx[1] + x[3][2] * xx[2];
  if (!(fabs(a) > CS_MACH_EPS)) {
    // the quadratic / linear branch has no transcendental: run it exactly as
    // isCoplanar does (same operations, so the verdict is exact)
    double r[3] = {-1.0, -1.0, -1.0};
    const double qa = b, qb = c, qc = d;
    const double delta = qb * qb - 4.0 * qa * qc;
    if (fabs(qa) > CS_ROUND_EPS && delta > 0) {
      const double ds = sqrt(delta);
      r[0] = (-qb + ds) / (2.0 * qa);
      r[1] = (-qb - ds) / (2.0 * qa);
    } else if (fabs(qa) < CS_ROUND_EPS && fabs(qb) > CS_ROUND_EPS) {
      r[0] = -qc / qb;
    }
    bool any = false;
    for (int i = 0; i < 3; ++i) {
      double t = r[i] - CS_MACH_EPS;
      if (t < 0 || t > dt) t = -1;
      any = any || t > CS_MACH_EPS;
    }
    return !any;
  }
  b /= a;
  c /= a;
  d /= a;
  a = b;
  b = c;
  c = d;
  const double Q = (a * a - 3 * b) / 9;
  const double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
  const double Q3 = Q * Q * Q, R2 = R * R;
  double r[3];
  int nr = 0;
  double scale = fabs(a / 3);
  if (R2 < Q3) {
    const double Qsqrt = sqrt(Q);
    const double theta = cs_fast_acos(R / sqrt(Q3));
    r[0] = -2 * Qsqrt * cs_fast_cos(theta / 3) - a / 3;
    r[1] = -2 * Qsqrt * cs_fast_cos((theta + 2 * CS_PI) / 3) - a / 3;
    r[2] = -2 * Qsqrt * cs_fast_cos((theta - 2 * CS_PI) / 3) - a / 3;
    nr = 3;
    scale += 2 * Qsqrt;
  } else {
    const double sgn = (R > 0) ? 1.0 : -1.0;
    const double A = -sgn * cs_cbrt_seed(fabs(R) + sqrt(R2 - Q3));
    const double mA = 1e-13 * fabs(A);
    if (!(fabs(fabs(A) - CS_ROUND_EPS) > mA)) return false;  // near the B = 0 threshold
    const double B = (fabs(A) < CS_ROUND_EPS) ? 0.0 : Q / A;
    scale += fabs(A) + fabs(B);
    r[0] = (A + B) - a / 3.0;
    nr = 1;
    if (!(fabs(fabs(A - B) - CS_ROUND_EPS) > 1e-13 * (fabs(A) + fabs(B))))
      return false;  // near the double-root threshold
    if (fabs(A - B) < CS_ROUND_EPS) {
      r[1] = -0.5 * (A + B) - a / 3.0;
      nr = 2;
    }
  }
  const double margin = 1e-13 * (scale + dt);
  for (int i = 0; i < nr; ++i) {
    const double t = r[i] - CS_MACH_EPS;
    // counts for isCoplanar iff MACH_EPS < t <= dt
    if (!(t + margin <= CS_MACH_EPS || t - margin > dt)) return false;
  }
  return true;
}

// isCoplanar (dcollid3d.cpp:371-482)
__device__ bool d_is_coplanar(const Sub& S, double dt, double roots[4]) {
  double v[4][3], x[4][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v[0][k] = S.v[0][k];
    x[0][k] = S.xo[0][k];
  }
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[i][k] = S.v[i][k] - S.v[0][k];
      x[i][k] = S.xo[i][k] - S.xo[0][k];
    }
  double a, b, c, d, vv[3], vx[3], xx[3];
  vv[0] = v[1][1] * v[2][2] - v[1][2] * v[2][1];
  vv[1] = v[1][0] * v[2][2] - v[1][2] * v[2][0];
  vv[2] = v[1][0] * v[2][1] - v[1][1] * v[2][0];
  vx[0] = v[1][1] * x[2][2] - v[1][2] * x[2][1] - v[2][1] * x[1][2] + v[2][2] * x[1][1];
  vx[1] = v[1][0] * x[2][2] - v[1][2] * x[2][0] - v[2][0] * x[1][2] + v[2][2] * x[1][0];
  vx[2] = v[1][0] * x[2][1] - v[1][1] * x[2][0] - v[2][0] * x[1][1] + v[2][1] * x[1][0];
  xx[0] = x[1][1] * x[2][2] - x[1][2] * x[2][1];
  xx[1] = x[1][0] * x[2][2] - x[1][2] * x[2][0];
  xx[2] = x[1][0] * x[2][1] - x[1][1] * x[2][0];
  a = v[3][0] * vv[0] - v[3][1] * vv[1] + v[3][2] * vv[2];
  b = x[3][0] * vv[0] - x[3][1] * vv[1] + x[3][2] * vv[2] + v[3][0] * vx[0] - v[3][1] * vx[1] +
      v[3][2] * vx[2];
  c = x[3][0] * vx[0] - x[3][1] * vx[1] + x[3][2] * vx[2] + v[3][0] * xx[0] - v[3][1] * xx[1] +
      v[3][2] * xx[2];
  d = x[3][0] * xx[0] - x[3][1] * xx[1] + x[3][2] * xx[2];
  roots[0] = roots[1] = roots[2] = -1.0;
  roots[3] = dt;
  if (fabs(a) > CS_MACH_EPS) {
    b /= a;
    c /= a;
    d /= a;
    a = b;
    b = c;
    c = d;
    const double Q = (a * a - 3 * b) / 9;
    const double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
    const double Q3 = Q * Q * Q, R2 = R * R;
    if (R2 < Q3) {
      const double Qsqrt = sqrt(Q);
      const double theta = cs_det_acos(R / sqrt(Q3));
      roots[0] = -2 * Qsqrt * cs_det_cos(theta / 3) - a / 3;
      roots[1] = -2 * Qsqrt * cs_det_cos((theta + 2 * CS_PI) / 3) - a / 3;
      roots[2] = -2 * Qsqrt * cs_det_cos((theta - 2 * CS_PI) / 3) - a / 3;
    } else {
      const double sgn = (R > 0) ? 1.0 : -1.0;
      const double A = -sgn * cs_det_pow_third(fabs(R) + sqrt(R2 - Q3));
      const double B = (fabs(A) < CS_ROUND_EPS) ? 0.0 : Q / A;
      roots[0] = (A + B) - a / 3.0;
      if (fabs(A - B) < CS_ROUND_EPS) roots[1] = roots[2] = -0.5 * (A + B) - a / 3.0;
    }
  } else {
    a = b;
    b = c;
    c = d;
    const double delta = b * b - 4.0 * a * c;
    if (fabs(a) > CS_ROUND_EPS && delta > 0) {
      const double ds = sqrt(delta);
      roots[0] = (-b + ds) / (2.0 * a);
      roots[1] = (-b - ds) / (2.0 * a);
    } else if (fabs(a) < CS_ROUND_EPS && fabs(b) > CS_ROUND_EPS) {
      roots[0] = -c / b;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    roots[i] = roots[i] - CS_MACH_EPS;
    if (roots[i] < 0 || roots[i] > dt) roots[i] = -1;
  }
  double t;
  if (roots[0] > roots[1]) { t = roots[0]; roots[0] = roots[1]; roots[1] = t; }
  if (roots[0] > roots[2]) { t = roots[0]; roots[0] = roots[2]; roots[2] = t; }
  if (roots[1] > roots[2]) { t = roots[1]; roots[1] = roots[2]; roots[2] = t; }
  return roots[0] > CS_MACH_EPS || roots[1] > CS_MACH_EPS || roots[2] > CS_MACH_EPS;
}

// roots / *coplanar: isCoplanar's outcome for a CCD sub-test (roots -1 and
// false when the certified screen decided it)
__device__ bool d_subtest(const Sub& S, bool is_pt, const NarrowParams& P, Contrib& C, int* err,
                          double roots[4], bool* coplanar, int known) {
  double c[4][3];
  if (!P.ccd) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) c[q][k] = S.xo[q][k];
    return is_pt ? d_point_to_tri(S, c, P.thickness, P, C, err)
                 : d_edge_to_edge(S, c, P.thickness, P, C, err);
  }
  roots[0] = roots[1] = roots[2] = -1.0;
  roots[3] = P.dt;
  *coplanar = false;
  // known: 1 = the certified screen already proved isCoplanar false, 0 = it
  // could not, -1 = not screened yet
  if (known == 1) return false;
  if (known < 0 && P.mixed && d_coplanar_certainly_false(S, P.dt)) return false;
  if (!d_is_coplanar(S, P.dt, roots)) return false;
  *coplanar = true;
  for (int r = 0; r < 4; ++r) {
    if (roots[r] < 0) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) c[q][k] = S.xo[q][k] + roots[r] * S.v[q][k];
    const bool hit = is_pt ? d_point_to_tri(S, c, P.round_eps, P, C, err)
                           : d_edge_to_edge(S, c, P.round_eps, P, C, err);
    if (hit) return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 screen of a static proximity sub-test.
//
// The points are taken relative to a local origin (point 2 for PointToTri,
// point 0 for EdgeToEdge, as the FP64 formulas do) in FP64 and rounded once
// to FP32, so every input carries a relative error of at most U = 2^-24
// (plus 2^-53).  Each decision quantity then gets a forward-error bound from
// the magnitudes of its operands (m_i = |x_i|):
//   dot product of two inputs      <= 5.1 U |a| |b|   (input rounding + 3 terms)
//   cross-product component        <= 4.1 U |a| |b|   (vector: <= 7.1 U |a| |b|)
//   det = d11 d22 - d12^2          <= 24 U m1^2 m2^2
//   N0, N1 (barycentric numerators)<= 24 U m1 m2 m3 max(m1, m2)
//   x43 . n, |n|                   <= 11.1 U m1 m2 m3,  9.1 U m1 m2
// and the code uses bounds with a safety factor of two to three above these.
// The screen reports a certain reject only when a quantity lies outside the
// FP64 test's acceptance region by more than its bound (and by a relative
// 1e-9 for the FP64 test's own rounding): then the FP64 test rejects too.
// Every other sub-test runs the FP64 test unchanged, so results are those of
// the FP64 path bit for bit.  Where the FP64 test would flag a degenerate
// normal (the reference's clean_up(ERROR)), the screen never rejects: it
// requires a certified non-degenerate triangle, and an edge pair's flag is
// raised only after its distance test has passed.
constexpr float MX_U = 5.9604645e-8f;  // 2^-24

struct MixedErr {
  float dist, param;  // FP32 - FP64 errors, for the statistics
  float ratio;        // largest |FP32 - FP64| / certified bound of a screened quantity
};
__device__ inline void mx_ratio(MixedErr* me, double err, float bound) {
  if (bound > 0.0f) me->ratio = fmaxf(me->ratio, (float)(err / (double)bound));
}

__device__ inline float f3dot(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline float f3mag(const float* a) { return sqrtf(f3dot(a, a)); }

// true: the FP64 PointToTri with thickness h and barycentric tolerance eps
// certainly rejects
__device__ bool screen_pt(const Sub& S, double h, double eps, bool stats, MixedErr* me) {
  float x13[3], x23[3], x43[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x13[k] = (float)(S.xo[0][k] - S.xo[2][k]);
    x23[k] = (float)(S.xo[1][k] - S.xo[2][k]);
    x43[k] = (float)(S.xo[3][k] - S.xo[2][k]);
  }
  const float m1 = f3mag(x13) * 1.000001f, m2 = f3mag(x23) * 1.000001f, m3 = f3mag(x43) * 1.000001f;
  const float d11 = f3dot(x13, x13), d22 = f3dot(x23, x23), d12 = f3dot(x13, x23);
  const float d13 = f3dot(x13, x43), d23 = f3dot(x23, x43);
  const float det = d11 * d22 - d12 * d12;
  const float e_det = 64.0f * MX_U * (m1 * m1) * (m2 * m2);
  if (!(det > 2.0f * e_det) || !(det > 1e-30f)) return false;  // not certainly well shaped
  const float n0 = d13 * d22 - d23 * d12, n1 = d11 * d23 - d12 * d13;
  const float mm = fmaxf(m1, m2);
  const float e_n = 64.0f * MX_U * m1 * m2 * m3 * mm;
  const float w0 = n0 / det, w1 = n1 / det, w2 = 1.0f - w0 - w1;
  const float den = det - e_det;
  const float ew0 = (e_n + fabsf(w0) * e_det) / den + 4.0f * MX_U * fabsf(w0);
  const float ew1 = (e_n + fabsf(w1) * e_det) / den + 4.0f * MX_U * fabsf(w1);
  const float ew2 = ew0 + ew1 + 4.0f * MX_U * (1.0f + fabsf(w0) + fabsf(w1));
  float n[3] = {x13[1] * x23[2] - x13[2] * x23[1], x13[2] * x23[0] - x13[0] * x23[2],
                x13[0] * x23[1] - x13[1] * x23[0]};
  const float nn = f3mag(n);
  const float e_nn = 28.0f * MX_U * m1 * m2;
  if (!(nn > 2.0f * e_nn)) return false;
  const float num = fabsf(f3dot(x43, n));
  const float dist = num / nn;
  const float e_dist = (32.0f * MX_U * m1 * m2 * m3 + dist * e_nn) / (nn - e_nn) + 4.0f * MX_U * dist;
  if (stats) {
    // the FP64 quantities of the same sub-test (PointToTri's formulas)
    double a[3], b[3], c[3], nd[3];
    sub3(S.xo[0], S.xo[2], a);
    sub3(S.xo[1], S.xo[2], b);
    sub3(S.xo[3], S.xo[2], c);
    const double D = dot3(a, a) * dot3(b, b) - dot3(a, b) * dot3(a, b);
    const double W0 = (dot3(a, c) * dot3(b, b) - dot3(b, c) * dot3(a, b)) / D;
    const double W1 = (dot3(a, a) * dot3(b, c) - dot3(a, b) * dot3(a, c)) / D;
    cross3(a, b, nd);
    const double dd = fabs(dot3(c, nd)) / mag3(nd);
    me->dist = fmaxf(me->dist, (float)(fabs((double)dist - dd) / h));
    me->param = fmaxf(me->param, (float)fmax(fabs((double)w0 - W0), fabs((double)w1 - W1)));
    mx_ratio(me, fabs((double)dist - dd), e_dist);
    mx_ratio(me, fabs((double)w0 - W0), ew0);
    mx_ratio(me, fabs((double)w1 - W1), ew1);
    mx_ratio(me, fabs((double)w2 - (1.0 - W0 - W1)), ew2);
  }
  const float slack = 1e-9f;
  if ((double)(dist - e_dist) > h * (1.0 + 1e-9)) return true;
  const float w[3] = {w0, w1, w2}, ew[3] = {ew0, ew1, ew2};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if ((double)(w[i] - ew[i]) > 1.0 + eps + slack) return true;
    if ((double)(w[i] + ew[i]) < -eps - slack) return true;
  }
  return false;
}

// true: the FP64 EdgeToEdge with thickness h certainly rejects.  The clamped
// segment distance is at least the gap between the segments' boxes and at
// least the distance between the two lines.
__device__ bool screen_ee(const Sub& S, double h, bool stats, MixedErr* me) {
  float p[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) p[q][k] = (float)(S.xo[q][k] - S.xo[0][k]);
  float gap = 0.0f, ext = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float alo = fminf(p[0][k], p[1][k]), ahi = fmaxf(p[0][k], p[1][k]);
    const float blo = fminf(p[2][k], p[3][k]), bhi = fmaxf(p[2][k], p[3][k]);
    gap = fmaxf(gap, fmaxf(blo - ahi, alo - bhi));
    ext = fmaxf(ext, fmaxf(fmaxf(fabsf(alo), fabsf(ahi)), fmaxf(fabsf(blo), fabsf(bhi))));
  }
  // a box gap: two subtractions of inputs of relative error U
  const float e_gap = 4.0f * MX_U * ext;
  if (stats) {  // the same gap from the FP64 differences
    double G = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double a0 = 0.0, a1 = S.xo[1][k] - S.xo[0][k];
      const double b0 = S.xo[2][k] - S.xo[0][k], b1 = S.xo[3][k] - S.xo[0][k];
      G = fmax(G, fmax(fmin(b0, b1) - fmax(a0, a1), fmin(a0, a1) - fmax(b0, b1)));
    }
    me->dist = fmaxf(me->dist, (float)(fabs((double)gap - G) / h));
    mx_ratio(me, fabs((double)gap - G), e_gap);
  }
  if ((double)(gap - e_gap) > h * (1.0 + 1e-9)) return true;
  float x21[3], x43[3], x31[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x21[k] = p[1][k];
    x43[k] = p[3][k] - p[2][k];
    x31[k] = p[2][k];
  }
  const float m21 = f3mag(x21) * 1.000001f, m43 = f3mag(x43) * 1.000001f, m31 = f3mag(x31) * 1.000001f;
  float n[3] = {x21[1] * x43[2] - x21[2] * x43[1], x21[2] * x43[0] - x21[0] * x43[2],
                x21[0] * x43[1] - x21[1] * x43[0]};
  const float nn = f3mag(n);
  // x43 = p3 - p2 carries one more subtraction: its error is at most 3 U m43 + 2 U ext
  const float e_nn = 40.0f * MX_U * (m21 * (m43 + ext));
  if (!(nn > 2.0f * e_nn)) return false;
  const float num = fabsf(f3dot(x31, n));
  const float lined = num / nn;
  const float e_num = 48.0f * MX_U * m31 * m21 * (m43 + ext);
  const float e_line = (e_num + lined * e_nn) / (nn - e_nn) + 4.0f * MX_U * lined;
  if (stats) {
    double a[3], b[3], c[3], nd[3];
    sub3(S.xo[1], S.xo[0], a);
    sub3(S.xo[3], S.xo[2], b);
    sub3(S.xo[2], S.xo[0], c);
    cross3(a, b, nd);
    const double L = fabs(dot3(c, nd)) / mag3(nd);
    me->dist = fmaxf(me->dist, (float)(fabs((double)lined - L) / h));
    mx_ratio(me, fabs((double)lined - L), e_line);
  }
  return (double)(lined - e_line) > h * (1.0 + 1e-9);
}

// local point indices of sub-test s.  kind: 0 tri-tri (A 0..2, B 3..5),
// 1 tri-bond (tri 0..2, bond 3..4), 2 bond-bond (A 0..1, B 2..3)
__device__ inline void subtest_points(int kind, int s, int ccd, int q[4], bool* is_pt) {
  if (kind == 0) {
    if (s < 6) {
      const int k = s / 3, i = s - 3 * k;
      // proximity: k=0 triangle B, point A[i]; CCD: k=0 triangle A, point B[i]
      const bool tri_is_a = (k == 0) == (ccd != 0);
      const int tb = tri_is_a ? 0 : 3, pb = tri_is_a ? 3 : 0;
      q[0] = tb; q[1] = tb + 1; q[2] = tb + 2; q[3] = pb + i;
      *is_pt = true;
    } else {
      const int i = (s - 6) / 3, j = (s - 6) - 3 * i;
      q[0] = i; q[1] = (i + 1) % 3; q[2] = 3 + j; q[3] = 3 + (j + 1) % 3;
      *is_pt = false;
    }
  } else if (kind == 1) {
    if (s < 2) {
      q[0] = 0; q[1] = 1; q[2] = 2; q[3] = 3 + s;
      *is_pt = true;
    } else {
      const int i = s - 2;
      q[0] = i; q[1] = (i + 1) % 3; q[2] = 3; q[3] = 4;
      *is_pt = false;
    }
  } else {
    q[0] = 0; q[1] = 1; q[2] = 2; q[3] = 3;
    *is_pt = false;
  }
}

__device__ inline int uf_load(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ int uf_find(int* uf, int x) {
  for (;;) {
    const int p = uf_load(&uf[x]);
    if (p == x) return x;
    const int gp = uf_load(&uf[p]);
    if (gp == p) return p;
    atomicCAS(&uf[x], p, gp);  // path halving; parents only decrease
    x = gp;
  }
}
__device__ void uf_union(int* uf, int a, int b) {
  for (;;) {
    a = uf_find(uf, a);
    b = uf_find(uf, b);
    if (a == b) return;
    if (a < b) { const int t = a; a = b; b = t; }
    if (atomicCAS(&uf[a], a, b) == a) return;  // larger root under smaller
  }
}

// zone membership: list each vertex once per step, the first time it is an
// endpoint of a union of two distinct vertices
struct ZoneList {
  int* uf;
  int* mark;
  uint32_t* list;
  int* count;
};
__device__ inline void zone_touch(const ZoneList& Z, int v) {
  if (__hip_atomic_load(&Z.mark[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  if (atomicExch(&Z.mark[v], 1) == 0) Z.list[atomicAdd(Z.count, 1)] = (uint32_t)v;
}
__device__ inline void zone_union(const ZoneList& Z, int a, int b) {
  if (a == b) return;
  zone_touch(Z, a);
  zone_touch(Z, b);
  uf_union(Z.uf, a, b);
}
// createImpZone(pts, 4, NO) (dcollid.cpp:473-484): every pair of the four
// points not holding a movable point is merged, which joins all non-movable
// points of the sub-test; they are merged through the first of them
__device__ inline void zone_merge4(const ZoneList& Z, const int32_t* vbody, const int p[4]) {
  bool ok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) ok[i] = !(vbody && vbody[p[i]] >= 0);
  int hub = -1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!ok[i]) continue;
    if (hub < 0) hub = p[i];
    else zone_union(Z, hub, p[i]);
  }
}

// vertices of an element pair (canonical orientation: A = hi, B = lo) and
// its kind: 0 tri-tri (15 sub-tests), 1 tri-bond (5), 2 bond-bond (1)
__device__ inline void pair_points(int2 pr, const int4* __restrict__ ev, int* A, int* B, int* kind,
                                   int* nsub, int g[6]) {
  *A = pr.y;
  *B = pr.x;
  const int4 ta = ev[*A], tb = ev[*B];
  const bool tri_a = ta.z >= 0, tri_b = tb.z >= 0;
  if (tri_a && tri_b) {
    *kind = 0; *nsub = 15;
    g[0] = ta.x; g[1] = ta.y; g[2] = ta.z; g[3] = tb.x; g[4] = tb.y; g[5] = tb.z;
  } else if (tri_a || tri_b) {
    const int4 T = tri_a ? ta : tb, Bd = tri_a ? tb : ta;
    *kind = 1; *nsub = 5;
    g[0] = T.x; g[1] = T.y; g[2] = T.z; g[3] = Bd.x; g[4] = Bd.y; g[5] = Bd.y;
  } else {
    *kind = 2; *nsub = 1;
    g[0] = ta.x; g[1] = ta.y; g[2] = tb.x; g[3] = tb.y; g[4] = tb.y; g[5] = tb.y;
  }
}

__device__ inline uint32_t subtest_hash(int B, int A, int s, uint32_t salt) {
  uint64_t z = ((uint64_t)(uint32_t)B << 32 | (uint32_t)A) ^ ((uint64_t)s << 58) ^
               ((uint64_t)salt * 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

// first root isCoplanar reports valid (the roots are sorted; -1 = invalid)
__device__ inline double first_valid_root(const double roots[4]) {
  for (int r = 0; r < 3; ++r)
    if (roots[r] >= 0) return roots[r];
  return -1.0;
}

// dataset sample of a CCD sub-test: features, root-exists label (isCoplanar)
// and the first valid root / dt
__device__ void collect_subtest(const Sub& S, const NarrowParams& P, int B, int A, int s,
                                bool coplanar, const double roots[4]) {
  if (subtest_hash(B, A, s, P.ds_salt) % (uint32_t)P.ds_stride != 0) return;
  const unsigned long long slot = atomicAdd(P.ds_count, 1ull);
  if ((long long)slot >= P.ds_cap) return;
  float f[MF_IN];
  subtest_features(S.xo, S.v, P.dt, P.thickness > 0 ? P.thickness : 1.0, f);
  for (int k = 0; k < MF_IN; ++k) P.dsx[slot * MF_IN + k] = f[k];
  const double r = coplanar ? first_valid_root(roots) : -1.0;
  P.dse[slot] = coplanar ? 1.f : 0.f;
  P.dst[slot] = (coplanar && r >= 0) ? (float)(r / P.dt) : 0.f;
}

__device__ inline int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

// Learned-form statistics of a scored CCD sub-test (every lane of the
// wavefront calls it, active = scored; one atomic per statistic and
// wavefront, into the work-group's copy): prediction vs isCoplanar, and for a sub-test with roots the
// bracketed Newton solve from the network's seed and from dt / 2, compared
// with the nearest valid closed-form root (+ MACH_EPS, which isCoplanar
// subtracts)
__device__ void learned_stats(const NarrowParams& P, bool active, bool coplanar,
                              const double roots[4], float sv, const Sub& S, int* sp) {
  bool pred = false;
  int found = 0, its = 0, itm = 0;
  float emax = 0.0f;
  if (active) {
    pred = signbit(sv);
    if (coplanar) {
      double a, b, c, d, rn = 0.0, rm = 0.0;
      cubic_coeffs(S.xo, S.v, &a, &b, &c, &d);
      // the window isCoplanar accepts (a root r with r - MACH_EPS in [0, dt])
      const int it = newton_first_root(a, b, c, d, CS_MACH_EPS, P.dt + CS_MACH_EPS,
                                       (double)fabsf(sv) * P.dt, &rn);
      const int im = newton_first_root(a, b, c, d, CS_MACH_EPS, P.dt + CS_MACH_EPS, 0.5 * P.dt, &rm);
      if (it >= 0) {
        double e = 1e300;
        for (int r = 0; r < 3; ++r)
          if (roots[r] >= 0) e = fmin(e, fabs(rn - (roots[r] + CS_MACH_EPS)));
        found = 1;
        its = it;
        itm = im >= 0 ? im : 0;
        if (e < 1e300) emax = (float)(e / P.dt);
      }
    }
  }
  const uint64_t b_sc = __ballot(active), b_ex = __ballot(active && coplanar);
  const uint64_t b_pr = __ballot(active && pred), b_ag = __ballot(active && pred == coplanar);
  const uint64_t b_nt = __ballot(found != 0);
  int sums[3] = {0, 0, 0};
  float em = 0.0f;
  if (b_nt) {
    sums[0] = wave_sum(found);
    sums[1] = wave_sum(its);
    sums[2] = wave_sum(itm);
    em = wave_max(emax);
  }
  if ((threadIdx.x & 63) == 0 && b_sc) {
    atomicAdd(&sp[SP_LN_SCORED], __popcll(b_sc));
    atomicAdd(&sp[SP_LN_EXISTS], __popcll(b_ex));
    atomicAdd(&sp[SP_LN_PRED], __popcll(b_pr));
    atomicAdd(&sp[SP_LN_AGREE], __popcll(b_ag));
    if (b_nt) {
      atomicAdd(&sp[SP_LN_ROOTS], sums[0]);
      atomicAdd(&sp[SP_LN_ITERS], sums[1]);
      atomicAdd(&sp[SP_LN_ITERS_MID], sums[2]);
      atomicMax(&sp[SP_LN_EMAX], __float_as_int(em));
    }
  }
}

// One group of 16 lanes per pair, lane s runs sub-test s (15 for two
// triangles, 5 for a triangle and a bond, 1 for two bonds), so the dependent
// chain of a lane is one sub-test.  The group's hit mask comes from one
// ballot; its first lane writes the hit-pair record, and every lane from the
// first hit onward performs its own zone merges (the cumulative-status
// rule).  Hit records take one atomic per wavefront; their order does not
// matter (they are sorted by key before use).

// One vertex of a pair as its sub-test lanes read it, staged in LDS.
struct StagedVertex {
  dvec4 xo, av;
  int fixed, body, pad0, pad1;
};

__device__ __attribute__((always_inline)) inline void narrow_lane(
    int64_t gl, const int2* __restrict__ pairs, int npairs, const int4* __restrict__ ev,
    const dvec4* __restrict__ x_old, const dvec4* __restrict__ av,
    const uint8_t* __restrict__ fixed, const NarrowParams& P, const ZoneList& Z,
    hit_rec* __restrict__ hits, int hit_cap, int4* __restrict__ hp, int hp_cap,
    int* __restrict__ counters, StagedVertex* __restrict__ stage) {
  int* err = counters + CNT_ERR;
  const int64_t i = gl / SUB_GROUP;
  const int s = (int)(gl % SUB_GROUP);
  const int lane = threadIdx.x & 63;
  int kind = 0, nsub = 0, A = 0, B = 0;
  int g[6] = {0, 0, 0, 0, 0, 0};
  int64_t pi = i;  // position of the pair in the stream
  if (i < npairs) {
    if (P.perm) pi = P.perm[i];
    pair_points(pairs[pi], ev, &A, &B, &kind, &nsub, g);
  }
  const bool active = s < nsub;
  int q[4] = {0, 0, 0, 0};
  bool hit = false, rejected = false, coplanar = false;
  MixedErr me{0.0f, 0.0f, 0.0f};
  int known = -1;     // learned form: the screen's verdict for this slot
  float lcode = 0.0f;
  if (P.ccd && P.lseed && active) {
    lcode = P.lseed[pi * SUB_GROUP + s];
    known = lcode == SEED_CERTAIN ? 1 : 0;
  }
  double roots[4] = {-1.0, -1.0, -1.0, -1.0};
  Sub S;
  Contrib C;
  // The pair's (up to) six vertices are loaded once per 16-lane group, lane s
  // loading vertex g[s], and every sub-test lane reads its four from LDS.
  // (The group is 16 lanes of one wavefront, whose LDS operations complete in
  // order: no barrier.)
  if (i < npairs && s < 6) {
    int vid = g[0];
#pragma unroll
    for (int k = 1; k < 6; ++k)
      if (s == k) vid = g[k];
    stage[s].xo = x_old[vid];
    stage[s].av = av[vid];
    stage[s].fixed = fixed[vid];
    stage[s].body = P.vbody ? P.vbody[vid] : -1;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (active) {
    bool is_pt;
    subtest_points(kind, s, P.ccd, q, &is_pt);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const StagedVertex& v = stage[q[p]];
      S.xo[p][0] = v.xo.x; S.xo[p][1] = v.xo.y; S.xo[p][2] = v.xo.z;
      S.v[p][0] = v.av.x; S.v[p][1] = v.av.y; S.v[p][2] = v.av.z;
      S.st[p] = v.fixed != 0;
      S.bd[p] = v.body;
      S.mv[p] = S.bd[p] >= 0;
      int vid = g[0];
#pragma unroll
      for (int k = 1; k < 6; ++k)
        if (q[p] == k) vid = g[k];
      q[p] = vid;  // from here on: global vertex ids
    }
    if (P.mixed && !P.ccd) {
      rejected = is_pt ? screen_pt(S, P.thickness, P.round_eps, P.mixed_stats, &me)
                       : screen_ee(S, P.thickness, P.mixed_stats, &me);
    }
    if (!rejected) hit = d_subtest(S, is_pt, P, C, err, roots, &coplanar, known);
    if (P.ccd && P.ds_cap > 0) collect_subtest(S, P, B, A, s, coplanar, roots);
  }
  int* sp = P.spread + (blockIdx.x & (SPREAD_SLOTS - 1)) * SPREAD_STRIDE;
  if (P.ccd && P.lseed) learned_stats(P, known == 0, coplanar, roots, lcode, S, sp);
  if (P.mixed && !P.ccd) {  // every lane of the wavefront takes part
    const uint64_t rb = __ballot(active && rejected), kb = __ballot(active && !rejected);
    float md = 0.0f, mp = 0.0f, mr = 0.0f;
    if (P.mixed_stats) {
      md = wave_max(me.dist);
      mp = wave_max(me.param);
      mr = wave_max(me.ratio);
    }
    if (lane == 0 && (rb | kb)) {
      atomicAdd(&sp[SP_MX_SCREENED], __popcll(rb));
      atomicAdd(&sp[SP_MX_RECHECK], __popcll(kb));
      if (P.mixed_stats) {
        atomicMax(&sp[SP_MX_EDIST], __float_as_int(md));
        atomicMax(&sp[SP_MX_EPARAM], __float_as_int(mp));
        atomicMax(&sp[SP_MX_RATIO], __float_as_int(mr));
      }
    }
  }
  // hit records: one atomic per wavefront
  const uint64_t hb = __ballot(hit);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int base = 0;
  if (lane == __ffsll((unsigned long long)hb) - 1) base = atomicAdd(&counters[CNT_HITS], __popcll(hb));
  base = __shfl(base, __ffsll((unsigned long long)hb) - 1);
  if (hit) {
    const int slot = base + __popcll(hb & lt);
    if (slot < hit_cap) {
      hit_rec& h = hits[slot];
      h.key = hit_key((uint32_t)B, (uint32_t)A, (uint32_t)s);
      h.sb[0] = C.sb[0];
      h.sb[1] = C.sb[1];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        h.v[p] = (S.st[p] || C.all_rigid) ? -1 : q[p];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          h.imp[p][k] = C.imp[p][k];
          h.fric[p][k] = C.fric[p][k];
        }
      }
    }
  }
  const int grp = lane / SUB_GROUP;
  const int mask = (int)((hb >> (grp * SUB_GROUP)) & 0xffffull);
  if (mask == 0) return;
  const int first_hit = __ffs(mask) - 1;
  if (s == 0) {
    const int slot = atomicAdd(&counters[CNT_HIT_PAIRS], 1);
    if (slot < hp_cap) hp[slot] = make_int4(B, A, mask, first_hit);
    atomicAdd(&counters[CNT_HIT_SUB], __popc(mask));
  }
  if (P.zone_mode && active && s >= first_hit) zone_merge4(Z, P.vbody, q);
}

// npairs_dev != nullptr: the pair count is read on the device (clamped to
// cap); the grid covers the first npairs pairs of it
template <int THREADS>
__global__ void __launch_bounds__(THREADS, 3)
    k_narrow(const int2* __restrict__ pairs, int npairs, const int* __restrict__ npairs_dev,
             int cap, const int4* __restrict__ ev, const dvec4* __restrict__ x_old,
             const dvec4* __restrict__ av, const uint8_t* __restrict__ fixed, NarrowParams P,
             ZoneList Z, hit_rec* __restrict__ hits, int hit_cap, int4* __restrict__ hp,
             int hp_cap, int* __restrict__ counters) {
  const int n = npairs_dev ? min(min(*npairs_dev, cap), npairs) : npairs;
  // XCD-contiguous shares of the pairs: the hardware deals work-groups to the
  // 8 XCDs round-robin, so the first `busy` of them (those the pair count
  // needs; with a count read on the device the grid may be larger) are spread
  // evenly, and remapping within them gives each XCD one contiguous share
  const int busy = (int)(((int64_t)n * SUB_GROUP + THREADS - 1) / THREADS);
  if ((int)blockIdx.x >= busy) return;
  const int b = xcd_block(blockIdx.x, busy);
  const int64_t gl = (int64_t)b * THREADS + threadIdx.x;
  if ((gl & ~63ll) >= (int64_t)n * SUB_GROUP) return;  // whole wavefronts past the end
  __shared__ StagedVertex stage[THREADS / SUB_GROUP][6];
  narrow_lane(gl, pairs, n, ev, x_old, av, fixed, P, Z, hits, hit_cap, hp, hp_cap, counters,
              stage[threadIdx.x / SUB_GROUP]);
}

static ZoneList zone_list(DevState& S) {
  return ZoneList{S.uf, S.zmark, S.zlist, S.counters + CNT_ZLIST};
}

// npairs_dev: the pair count is read on the device and only its first
// npairs pairs are covered (the caller checks the count afterwards)
void launch_narrow(hipStream_t st, DevState& S, const NarrowParams& P, const int2* pairs,
                   int npairs, const int* npairs_dev, int cap) {
  if (npairs <= 0) return;
  const int64_t lanes = (int64_t)npairs * SUB_GROUP;
  const int T = S.tune_narrow;
  const int64_t blocks = (lanes + T - 1) / T;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(T), 0, st, pairs, npairs, npairs_dev,
                       cap, S.ev, S.x_old, S.av, S.fixed, P, zone_list(S), S.hits, S.hit_cap, S.hp,
                       S.hp_cap, S.counters);
  };
  if (T == 64) go(k_narrow<64>);
  else if (T == 256) go(k_narrow<256>);
  else go(k_narrow<128>);
}

// Learned form, step 1: the certified coplanarity screen of every CCD
// sub-test of the pair stream (4 slots per thread, 16 per pair).  A slot's
// code in seeds[]: SEED_INACTIVE (no such sub-test), SEED_CERTAIN (the
// screen proves isCoplanar false) or 0 (to be scored; the network's output
// replaces it).  The uncertain slots are appended to `list` (count[0]) and
// the pairs holding one are counted (count[1]), one atomic per work-group
// each; every pair gets the key that sorts it after all scored pairs until
// a scored slot lowers it.
constexpr int SCREEN_THREADS = 256, SCREEN_SLOTS = 4 * SCREEN_THREADS;

__global__ void __launch_bounds__(SCREEN_THREADS)
    k_learned_screen(const int2* __restrict__ pairs, int npairs, const int* __restrict__ npairs_dev,
                     int cap, const int4* __restrict__ ev, const dvec4* __restrict__ x_old,
                     const dvec4* __restrict__ av, double dt, float* __restrict__ seeds,
                     uint32_t* __restrict__ keys, uint32_t* __restrict__ idx,
                     uint32_t* __restrict__ list, int* __restrict__ count) {
  __shared__ int wcount[SCREEN_THREADS / 64][2];
  __shared__ int wbase[SCREEN_THREADS / 64];
  const int n = npairs_dev ? min(min(*npairs_dev, cap), npairs) : npairs;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * SCREEN_SLOTS;
  bool unc[4];
  int mine = 0, mpairs = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // a pair's 16 slots are 16 consecutive lanes of one wavefront
    const int64_t slot = base + k * SCREEN_THREADS + t;
    const int64_t i = slot / SUB_GROUP;
    const int s = (int)(slot % SUB_GROUP);
    unc[k] = false;
    if (i < npairs) {
      float code = SEED_INACTIVE;
      if (i < n) {
        int A, B, kind, nsub, g[6];
        pair_points(pairs[i], ev, &A, &B, &kind, &nsub, g);
        if (s == 0) {
          keys[i] = KEY_NONE;
          idx[i] = (uint32_t)i;
        }
        if (s < nsub) {
          int q[4];
          bool is_pt;
          subtest_points(kind, s, 1, q, &is_pt);
          Sub S;
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            int vid = g[0];
#pragma unroll
            for (int m = 1; m < 6; ++m)
              if (q[p] == m) vid = g[m];
            const dvec4 a = x_old[vid], b = av[vid];
            S.xo[p][0] = a.x; S.xo[p][1] = a.y; S.xo[p][2] = a.z;
            S.v[p][0] = b.x; S.v[p][1] = b.y; S.v[p][2] = b.z;
          }
          unc[k] = !d_coplanar_certainly_false(S, dt);
          code = unc[k] ? 0.0f : SEED_CERTAIN;
        }
      } else if (s == 0) {
        keys[i] = KEY_NONE;    // past the device-side count: the stable sort keeps
        idx[i] = (uint32_t)i;  // it after every counted pair
      }
      seeds[slot] = code;
    }
    mine += unc[k];
    // pairs with at least one slot to score (counted by their slot 0)
    const uint64_t m = __ballot(unc[k]);
    if ((lane & (SUB_GROUP - 1)) == 0 && ((m >> (lane & ~(SUB_GROUP - 1))) & 0xffffull)) ++mpairs;
  }
  // one atomic per work-group and counter
  int wsum = mine, psum = mpairs;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    wsum += __shfl_xor(wsum, off);
    psum += __shfl_xor(psum, off);
  }
  int pre = mine;  // inclusive prefix within the wavefront
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(pre, off);
    if (lane >= off) pre += v;
  }
  if (lane == 0) {
    wcount[wave][0] = wsum;
    wcount[wave][1] = psum;
  }
  __syncthreads();
  if (t == 0) {
    int tot = 0, ptot = 0;
    for (int w = 0; w < SCREEN_THREADS / 64; ++w) {
      wbase[w] = tot;
      tot += wcount[w][0];
      ptot += wcount[w][1];
    }
    const int b0 = tot ? atomicAdd(&count[0], tot) : 0;
    if (ptot) atomicAdd(&count[1], ptot);
    for (int w = 0; w < SCREEN_THREADS / 64; ++w) wbase[w] += b0;
  }
  __syncthreads();
  int pos = wbase[wave] + pre - mine;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (unc[k]) list[pos++] = (uint32_t)(base + k * SCREEN_THREADS + t);
}

// Learned form, step 2: the network over the listed (uncertain) slots, 64
// per tile; each slot's seed sigmoid(s) in [0, 1] with the sign bit set when
// the logit predicts a root, and its pair's key lowered to the bucket of the
// best score (254 - floor(254 sigmoid(logit))).
__global__ void __launch_bounds__(MLP_THREADS)
    k_learned_mlp(const int2* __restrict__ pairs, const int* __restrict__ count,
                  const uint32_t* __restrict__ list, const int4* __restrict__ ev,
                  const dvec4* __restrict__ x_old, const dvec4* __restrict__ av, double dt,
                  double h, const uint16_t* __restrict__ wb, const float* __restrict__ w,
                  float* __restrict__ seeds, uint32_t* __restrict__ keys) {
  __shared__ __align__(16) uint16_t x[MT * MF_INP];
  __shared__ __align__(16) uint16_t h1[MT * MF_H];
  __shared__ __align__(16) uint16_t h2[MT * MF_H];
  __shared__ float o[MT * MF_OUT];
  const int nl = *count;
  const int ntile = (nl + MT - 1) / MT;
  const int t = threadIdx.x;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int e = tile * MT + t;
    uint32_t slot = 0;
    if (t < MT) {
      float f[MF_IN];
      const bool act = e < nl;
      if (act) {
        slot = list[e];
        const int i = (int)(slot / SUB_GROUP), s = (int)(slot % SUB_GROUP);
        int A, B, kind, nsub, g[6];
        pair_points(pairs[i], ev, &A, &B, &kind, &nsub, g);
        int q[4];
        bool is_pt;
        subtest_points(kind, s, 1, q, &is_pt);
        double xo[4][3], v[4][3];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          int vid = g[0];
#pragma unroll
          for (int m = 1; m < 6; ++m)
            if (q[p] == m) vid = g[m];
          const dvec4 a = x_old[vid], b = av[vid];
          xo[p][0] = a.x; xo[p][1] = a.y; xo[p][2] = a.z;
          v[p][0] = b.x; v[p][1] = b.y; v[p][2] = b.z;
        }
        subtest_features(xo, v, dt, h, f);
      }
#pragma unroll
      for (int k = 0; k < MF_INP; ++k)
        x[t * MF_INP + k] = (act && k < MF_IN) ? f2bf(f[k]) : (uint16_t)0;
    }
    __syncthreads();
    tile_forward(x, h1, h2, o, wb, w);
    if (t < MT && e < nl) {
      const float z = o[t * MF_OUT];
      const float sd = mlp_sigmoid(o[t * MF_OUT + 1]);
      seeds[slot] = z > 0.f ? -sd : sd;
      const float p = mlp_sigmoid(z);
      const int bucket = 254 - (int)fminf(fmaxf(floorf(p * 254.0f), 0.0f), 254.0f);
      atomicMin(&keys[slot / SUB_GROUP], (uint32_t)bucket);
    }
    __syncthreads();  // the tile's LDS is rewritten by the next one
  }
}

// Scores the stream and sorts its pairs by score bucket (one stable pass of
// the radix sort over 8-bit keys); *perm = the pair positions in test order,
// the pairs with a scored slot first (L.count[1] of them).
hipError_t launch_learned_score(hipStream_t st, DevState& S, LearnedState& L, const int2* pairs,
                                int npairs, const int* npairs_dev, int cap, double dt, double h,
                                const uint32_t** perm) {
  *perm = nullptr;
  if (npairs <= 0) return hipSuccess;
  if (npairs > L.score_cap) {
    const int c = npairs + npairs / 4 + 1024;
    void* bufs[] = {L.seeds, L.keys, L.idx, L.tkeys, L.tidx, L.hist, L.part, L.list};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    L.seeds = nullptr;
    L.keys = L.idx = L.tkeys = L.tidx = L.hist = L.part = L.list = nullptr;
    L.score_cap = 0;
    const size_t hw = radix_hist_words(c), pw = scan_partial_words((int)hw);
    const size_t slots = (size_t)(c + 64) * SUB_GROUP;  // whole screen blocks
    CS_HIP_RET(hipMalloc((void**)&L.seeds, sizeof(float) * slots));
    CS_HIP_RET(hipMalloc((void**)&L.list, sizeof(uint32_t) * slots));
    CS_HIP_RET(hipMalloc((void**)&L.keys, sizeof(uint32_t) * (size_t)c));
    CS_HIP_RET(hipMalloc((void**)&L.idx, sizeof(uint32_t) * (size_t)c));
    CS_HIP_RET(hipMalloc((void**)&L.tkeys, sizeof(uint32_t) * (size_t)c));
    CS_HIP_RET(hipMalloc((void**)&L.tidx, sizeof(uint32_t) * (size_t)c));
    CS_HIP_RET(hipMalloc((void**)&L.hist, sizeof(uint32_t) * hw));
    CS_HIP_RET(hipMalloc((void**)&L.part, sizeof(uint32_t) * pw));
    L.score_cap = c;
  }
  if (!L.count) CS_HIP_RET(hipMalloc((void**)&L.count, 2 * sizeof(int)));
  CS_HIP_RET(hipMemsetAsync(L.count, 0, 2 * sizeof(int), st));
  const int64_t nslots = (int64_t)npairs * SUB_GROUP;
  const int nblk = (int)((nslots + SCREEN_SLOTS - 1) / SCREEN_SLOTS);
  hipLaunchKernelGGL(k_learned_screen, dim3((unsigned)nblk), dim3(SCREEN_THREADS), 0, st, pairs,
                     npairs, npairs_dev, cap, S.ev, S.x_old, S.av, dt, L.seeds, L.keys, L.idx,
                     L.list, L.count);
  CS_HIP_RET(hipGetLastError());
  const int64_t max_tiles = (nslots + MT - 1) / MT;
  hipLaunchKernelGGL(k_learned_mlp, dim3((unsigned)std::min<int64_t>(max_tiles, 1024)),
                     dim3(MLP_THREADS), 0, st, pairs, L.count, L.list, S.ev, S.x_old, S.av, dt, h,
                     L.wb, L.w, L.seeds, L.keys);
  CS_HIP_RET(hipGetLastError());
  bool in_tmp = false;
  CS_HIP_RET(radix_sort_u32(st, L.keys, L.idx, L.tkeys, L.tidx, L.hist, L.part, npairs, 8, &in_tmp));
  *perm = in_tmp ? L.tidx : L.idx;
  return hipSuccess;
}

// impact-zone merges of hit pairs carried from the previous pass: the points
// of every sub-test from the pair's first hit onward
__global__ void k_hp_unions(const int4* __restrict__ hp, int nmax, const int* __restrict__ n_dev,
                            const int4* __restrict__ ev, int ccd, const int32_t* __restrict__ vbody,
                            ZoneList Z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= min(nmax, *n_dev)) return;
  const int4 r = hp[i];
  const int A = r.y, B = r.x;
  const int4 ta = ev[A], tb = ev[B];
  const bool tri_a = ta.z >= 0, tri_b = tb.z >= 0;
  int kind, nsub, g[6];
  if (tri_a && tri_b) {
    kind = 0; nsub = 15;
    g[0] = ta.x; g[1] = ta.y; g[2] = ta.z; g[3] = tb.x; g[4] = tb.y; g[5] = tb.z;
  } else if (tri_a || tri_b) {
    const int4 T = tri_a ? ta : tb, Bd = tri_a ? tb : ta;
    kind = 1; nsub = 5;
    g[0] = T.x; g[1] = T.y; g[2] = T.z; g[3] = Bd.x; g[4] = Bd.y; g[5] = Bd.y;
  } else {
    kind = 2; nsub = 1;
    g[0] = ta.x; g[1] = ta.y; g[2] = tb.x; g[3] = tb.y; g[4] = tb.y; g[5] = tb.y;
  }
  for (int s = r.w; s < nsub; ++s) {
    int q[4];
    bool is_pt;
    subtest_points(kind, s, ccd, q, &is_pt);
    const int p4[4] = {g[q[0]], g[q[1]], g[q[2]], g[q[3]]};
    zone_merge4(Z, vbody, p4);
  }
}

// the records carried into this pass: their count is the hit-pair counter
// before the narrow phase appends (n is an upper bound)
void launch_hp_unions(hipStream_t st, DevState& S, int n, int ccd) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hp_unions, dim3((n + 255) / 256), dim3(256), 0, st, S.hp, n,
                     S.counters + CNT_HIT_PAIRS, S.ev, ccd, S.vbody, zone_list(S));
}

// parity log: (pass, lo, hi, mask) for every hit pair of the pass
__global__ void k_hp_log(const int4* __restrict__ hp, int n, int pass_id, int4* __restrict__ plog,
                         int cap, int* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 r = hp[i];
  const int slot = atomicAdd(&counters[CNT_LOG], 1);
  if (slot < cap) plog[slot] = make_int4(pass_id, r.x, r.y, r.z);
}

void launch_hp_log(hipStream_t st, DevState& S, int n, int pass_id) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hp_log, dim3((n + 255) / 256), dim3(256), 0, st, S.hp, n, pass_id, S.plog,
                     S.plog_cap, S.counters);
}

__global__ void k_math_eval(int fn, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  y[i] = fn == 0 ? cs_det_acos(v) : fn == 1 ? cs_det_cos(v) : fn == 2 ? cs_det_sin(v)
                                                                   : cs_det_pow_third(v);
}

hipError_t launch_math_eval(hipStream_t st, int fn, const double* x, double* y, int64_t n) {
  hipLaunchKernelGGL(k_math_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fn, x, y, n);
  CS_HIP_RET(hipGetLastError());
  return hipStreamSynchronize(st);
}

}  // namespace cs
