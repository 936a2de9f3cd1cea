// Adapter from a FronTier-style interface to the flat mesh of the C ABI:
// what CollisionSolver3d::assembleFromInterface (dcollid3d.cpp:12-52) walks,
// in the same order.  FronTier itself is not a dependency: the caller
// supplies an access policy `A` that reads its interface objects, so a
// FronTier application instantiates it with accessors over INTERFACE,
// SURFACE, CURVE, TRI, BOND, POINT and STATE.
//
// Order (the reference's hseList): the triangles of every non-boundary
// surface (intfc_surface_loop, surf_tri_loop), then the bonds of every
// STRING_HSBDRY curve (intfc_curve_loop, curve_bond_loop).  Points are
// numbered once, in order of first appearance in that walk.
//
// The policy A provides (static member functions; F is any callable):
//   using Interface, Surface, Curve, Tri, Bond, Point;
//   void for_each_surface(const Interface*, F)       // intfc_surface_loop
//   bool is_bdry(const Surface*)
//   void for_each_tri(const Surface*, F)             // surf_tri_loop
//   Point* tri_point(const Tri*, int)                // Point_of_tri
//   double side_length0(const Tri*, int)             // rest length of edge (i, i+1)
//   void for_each_curve(const Interface*, F)         // intfc_curve_loop
//   bool is_string(const Curve*)                     // hsbdry_type == STRING_HSBDRY
//   void for_each_bond(const Curve*, F)              // curve_bond_loop
//   Point* bond_start(const Bond*), Point* bond_end(const Bond*)
//   double length0(const Bond*)
//   bool is_fixed(const Point*)                      // STATE.is_fixed
//   bool is_movable(const Point*)                    // STATE.is_movableRG
//   double body_mass(const Surface*)                 // total_mass of a movable surface
//   void body_com(const Surface*, double[3])         // center_of_mass of it
//   void get_coords(const Point*, double[3]), void set_coords(Point*, const double[3])
//   void get_vel(const Point*, double[3]),    void set_vel(Point*, const double[3])
#pragma once

#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../collision_amd.h"

namespace collision_amd {

template <class A>
class FronTierMesh {
 public:
  using Interface = typename A::Interface;
  using Surface = typename A::Surface;
  using Curve = typename A::Curve;
  using Tri = typename A::Tri;
  using Bond = typename A::Bond;
  using Point = typename A::Point;

  // assembleFromInterface's walk: element arrays, point numbering, bodies
  void build(const Interface* intfc) {
    points.clear();
    id_.clear();
    tris.clear();
    tri_surf.clear();
    tri_len0.clear();
    bonds.clear();
    bond_len0.clear();
    is_fixed.clear();
    vert_body.clear();
    body_mass.clear();
    body_com.clear();
    int surf_index = 0;
    A::for_each_surface(intfc, [&](const Surface* s) {
      if (A::is_bdry(s)) return;
      int body = -1;
      A::for_each_tri(s, [&](const Tri* t) {
        for (int i = 0; i < 3; ++i) {
          Point* p = A::tri_point(t, i);
          const int v = number(p);
          tris.push_back(v);
          if (A::is_movable(p) && body < 0) {
            body = (int)body_mass.size();
            body_mass.push_back(A::body_mass(s));
            double c[3];
            A::body_com(s, c);
            body_com.insert(body_com.end(), c, c + 3);
          }
          if (A::is_movable(p)) vert_body[(size_t)v] = body;
        }
        tri_surf.push_back(surf_index);
        for (int j = 0; j < 3; ++j) tri_len0.push_back(A::side_length0(t, j));
      });
      ++surf_index;
    });
    A::for_each_curve(intfc, [&](const Curve* c) {
      if (!A::is_string(c)) return;
      A::for_each_bond(c, [&](const Bond* b) {
        bonds.push_back(number(A::bond_start(b)));
        bonds.push_back(number(A::bond_end(b)));
        bond_len0.push_back(A::length0(b));
      });
    });
  }

  // the flat view of the arrays (valid while this object lives unchanged)
  cs_mesh_view view() const {
    cs_mesh_view m{};
    m.nv = (int32_t)points.size();
    m.nt = (int32_t)(tris.size() / 3);
    m.nb = (int32_t)(bonds.size() / 2);
    m.tris = tris.data();
    m.tri_surf = tri_surf.data();
    m.tri_len0 = tri_len0.data();
    m.bonds = bonds.data();
    m.bond_len0 = bond_len0.data();
    m.is_fixed = is_fixed.data();
    m.nbody = (int32_t)body_mass.size();
    m.vert_body = m.nbody ? vert_body.data() : nullptr;
    m.body_mass = m.nbody ? body_mass.data() : nullptr;
    m.body_com = m.nbody ? body_com.data() : nullptr;
    return m;
  }

  // Coords and STATE.vel of every numbered point, [nv][3] each
  void gather(std::vector<double>& x, std::vector<double>& vel) const {
    x.resize(3 * points.size());
    vel.resize(3 * points.size());
    for (size_t v = 0; v < points.size(); ++v) {
      A::get_coords(points[v], &x[3 * v]);
      A::get_vel(points[v], &vel[3 * v]);
    }
  }

  // write the step's results back into the points (resolveCollision mutates
  // Coords and STATE.vel in place)
  void scatter(const double* x, const double* vel) const {
    for (size_t v = 0; v < points.size(); ++v) {
      A::set_coords(points[v], x + 3 * v);
      A::set_vel(points[v], vel + 3 * v);
    }
  }

  std::vector<Point*> points;  // vertex id -> point
  std::vector<int32_t> tris, tri_surf, bonds, vert_body;
  std::vector<double> tri_len0, bond_len0, body_mass, body_com;
  std::vector<uint8_t> is_fixed;

 private:
  int number(Point* p) {
    auto it = id_.find(p);
    if (it != id_.end()) return it->second;
    const int v = (int)points.size();
    id_.emplace(p, v);
    points.push_back(p);
    is_fixed.push_back(A::is_fixed(p) ? 1 : 0);
    vert_body.push_back(-1);
    return v;
  }
  std::unordered_map<const Point*, int> id_;
};

}  // namespace collision_amd
