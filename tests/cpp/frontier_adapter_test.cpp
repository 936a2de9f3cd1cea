// Host test of include/collision_amd/FronTierAdapter.hpp over a small
// FronTier-like object model written for this test (surfaces of triangles
// over shared points with per-point state, curves of bonds; boundary
// surfaces and non-string curves are skipped by the walk).  Prints "ok" or
// the first failed check.
#include <cmath>
#include <cstdio>
#include <vector>

#include "collision_amd/FronTierAdapter.hpp"

namespace mock {
struct State {
  double vel[3] = {0, 0, 0};
  bool is_fixed = false, is_movable = false;
};
struct Point {
  double x[3];
  State st;
};
struct Tri {
  Point* p[3];
  double len0[3];
};
struct Surface {
  bool bdry = false;
  std::vector<Tri> tris;
  double mass = 0.0, com[3] = {0, 0, 0};
};
struct Bond {
  Point *start, *end;
  double len0;
};
struct Curve {
  bool string = false;
  std::vector<Bond> bonds;
};
struct Interface {
  std::vector<Surface*> surfaces;
  std::vector<Curve*> curves;
};

struct Access {
  using Interface = mock::Interface;
  using Surface = mock::Surface;
  using Curve = mock::Curve;
  using Tri = mock::Tri;
  using Bond = mock::Bond;
  using Point = mock::Point;
  template <class F> static void for_each_surface(const Interface* i, F f) {
    for (const Surface* s : i->surfaces) f(s);
  }
  static bool is_bdry(const Surface* s) { return s->bdry; }
  template <class F> static void for_each_tri(const Surface* s, F f) {
    for (const Tri& t : s->tris) f(&t);
  }
  static Point* tri_point(const Tri* t, int i) { return t->p[i]; }
  static double side_length0(const Tri* t, int j) { return t->len0[j]; }
  template <class F> static void for_each_curve(const Interface* i, F f) {
    for (const Curve* c : i->curves) f(c);
  }
  static bool is_string(const Curve* c) { return c->string; }
  template <class F> static void for_each_bond(const Curve* c, F f) {
    for (const Bond& b : c->bonds) f(&b);
  }
  static Point* bond_start(const Bond* b) { return b->start; }
  static Point* bond_end(const Bond* b) { return b->end; }
  static double length0(const Bond* b) { return b->len0; }
  static bool is_fixed(const Point* p) { return p->st.is_fixed; }
  static bool is_movable(const Point* p) { return p->st.is_movable; }
  static double body_mass(const Surface* s) { return s->mass; }
  static void body_com(const Surface* s, double c[3]) {
    for (int k = 0; k < 3; ++k) c[k] = s->com[k];
  }
  static void get_coords(const Point* p, double x[3]) {
    for (int k = 0; k < 3; ++k) x[k] = p->x[k];
  }
  static void set_coords(Point* p, const double x[3]) {
    for (int k = 0; k < 3; ++k) p->x[k] = x[k];
  }
  static void get_vel(const Point* p, double v[3]) {
    for (int k = 0; k < 3; ++k) v[k] = p->st.vel[k];
  }
  static void set_vel(Point* p, const double v[3]) {
    for (int k = 0; k < 3; ++k) p->st.vel[k] = v[k];
  }
};
}  // namespace mock

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);      \
      ++fails;                                                     \
    }                                                              \
  } while (0)

int main() {
  using namespace mock;
  // points: a square cloth (0..3), a fixed triangle (4..6), a movable
  // triangle (7..9), a string (10..12) hanging from cloth point 2, a
  // boundary surface point (13) and an unused curve point (14)
  std::vector<Point> P(15);
  for (int i = 0; i < 15; ++i) P[i] = Point{{(double)i, 0.5 * i, -0.25 * i}, State{}};
  for (int i = 4; i < 7; ++i) P[i].st.is_fixed = true;
  for (int i = 7; i < 10; ++i) P[i].st.is_movable = true;
  P[3].st.vel[2] = -1.0;
  Surface bdry;  // skipped
  bdry.bdry = true;
  bdry.tris.push_back(Tri{{&P[13], &P[0], &P[1]}, {9, 9, 9}});
  Surface cloth;
  cloth.tris.push_back(Tri{{&P[0], &P[1], &P[3]}, {1, 2, 3}});
  cloth.tris.push_back(Tri{{&P[0], &P[3], &P[2]}, {4, 5, 6}});
  Surface box;
  box.tris.push_back(Tri{{&P[6], &P[5], &P[4]}, {7, 8, 9}});
  Surface body;
  body.mass = 2.5;
  body.com[0] = 8.0; body.com[1] = 4.0; body.com[2] = -2.0;
  body.tris.push_back(Tri{{&P[7], &P[8], &P[9]}, {10, 11, 12}});
  Curve string;
  string.string = true;
  string.bonds.push_back(Bond{&P[2], &P[10], 0.1});
  string.bonds.push_back(Bond{&P[10], &P[11], 0.2});
  string.bonds.push_back(Bond{&P[11], &P[12], 0.3});
  Curve other;  // not a string: skipped
  other.bonds.push_back(Bond{&P[14], &P[12], 0.4});
  Interface intfc;
  intfc.surfaces = {&bdry, &cloth, &box, &body};
  intfc.curves = {&other, &string};

  collision_amd::FronTierMesh<Access> m;
  m.build(&intfc);
  const cs_mesh_view v = m.view();
  CHECK(v.nv == 13 && v.nt == 4 && v.nb == 3);
  // first-appearance numbering: 0,1,3,2 | 6,5,4 | 7,8,9 | 10,11,12
  const Point* order[13] = {&P[0], &P[1], &P[3], &P[2], &P[6], &P[5], &P[4],
                            &P[7], &P[8], &P[9], &P[10], &P[11], &P[12]};
  for (int i = 0; i < 13 && i < v.nv; ++i) CHECK(m.points[i] == order[i]);
  const int tris[12] = {0, 1, 2, 0, 2, 3, 4, 5, 6, 7, 8, 9};
  for (int i = 0; i < 12; ++i) CHECK(v.tris[i] == tris[i]);
  const int surf[4] = {0, 0, 1, 2};  // boundary surfaces are not numbered
  for (int i = 0; i < 4; ++i) CHECK(v.tri_surf[i] == surf[i]);
  for (int i = 0; i < 12; ++i) CHECK(v.tri_len0[i] == i + 1);
  const int bonds[6] = {3, 10, 10, 11, 11, 12};
  for (int i = 0; i < 6; ++i) CHECK(v.bonds[i] == bonds[i]);
  CHECK(v.bond_len0[0] == 0.1 && v.bond_len0[2] == 0.3);
  for (int i = 0; i < 13; ++i) CHECK(v.is_fixed[i] == (i >= 4 && i < 7 ? 1 : 0));
  CHECK(v.nbody == 1 && v.body_mass[0] == 2.5 && v.body_com[0] == 8.0 && v.body_com[2] == -2.0);
  for (int i = 0; i < 13; ++i) CHECK(v.vert_body[i] == (i >= 7 && i < 10 ? 0 : -1));
  // state round trip
  std::vector<double> x, vel;
  m.gather(x, vel);
  CHECK(x.size() == 39 && x[3 * 2 + 0] == 3.0 && vel[3 * 2 + 2] == -1.0);  // vertex 2 = P[3]
  for (double& d : x) d += 1.0;
  for (double& d : vel) d -= 2.0;
  m.scatter(x.data(), vel.data());
  CHECK(P[3].x[0] == 4.0 && P[3].st.vel[2] == -3.0 && P[13].x[0] == 13.0 && P[14].st.vel[0] == 0.0);
  if (fails == 0) std::printf("ok\n");
  return fails ? 1 : 0;
}
