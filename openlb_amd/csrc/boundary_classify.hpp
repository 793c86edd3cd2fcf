// Boundary type and discrete normal of a wall cell, host side.
//
// Restates OpenLB's computeBoundaryTypeAndNormal (src/geometry/
// discreteNormals.hh:174-1190), which boundary::set uses to pick each wall
// cell's local boundary dynamics (src/boundary/setBoundary.h:83-101).  The
// reference tests a fixed sequence of neighbour patterns -- flat faces,
// external corners, internal corners, external edges, internal edges -- and
// keeps the first match of the first four groups; an internal-edge match
// replaces whatever was found before.  The patterns are kept here as data:
// each token is a predicate on the neighbour at the given offset,
//   W  wall: neither fluid (id 1) nor outside (id 0)
//   F  fluid
//   N  not fluid
// and the normal points out of the fluid.  Axes flagged periodic wrap; beyond
// a non-periodic face lies the outside, or, for a block of a split lattice,
// the first overlap layer with the materials of the neighbouring blocks (the
// reference classifies on the block geometry, whose overlap holds them).
#pragma once

#include <cstdint>
#include <cstring>

namespace lbmh {

enum BoundaryType : int { kTypeFlat = 0, kTypeExternalCorner = 1, kTypeInternalCorner = 2, kTypeExternalEdge = 3,
                          kTypeInternalEdge = 4 };

struct NormalRule {
  int type;
  int normal[3];
  bool replaces;
  const char* pattern;
};

inline const NormalRule* normal_rules(int& count)
{
  static const NormalRule rules[] = {
    {0, {-1, 0, 0}, false, "W+0+0+1 W+0+0-1 W+0+1+0 W+0-1+0 F+1+0+0"},
    {0, {1, 0, 0}, false, "W+0+0+1 W+0+0-1 W+0+1+0 W+0-1+0 F-1+0+0"},
    {0, {0, -1, 0}, false, "W+0+0+1 W+0+0-1 W+1+0+0 W-1+0+0 F+0+1+0"},
    {0, {0, 1, 0}, false, "W+0+0+1 W+0+0-1 W+1+0+0 W-1+0+0 F+0-1+0"},
    {0, {0, 0, -1}, false, "W+1+0+0 W-1+0+0 W+0+1+0 W+0-1+0 F+0+0+1"},
    {0, {0, 0, 1}, false, "W+1+0+0 W-1+0+0 W+0+1+0 W+0-1+0 F+0+0-1"},
    {1, {-1, -1, -1}, false, "W+1+0+0 W+0+0+1 W+1+0+1 W+0+1+0 W+1+1+0 W+0+1+1 F+1+1+1"},
    {1, {-1, 1, -1}, false, "W+1+0+0 W+0+0+1 W+1+0+1 W+0-1+0 W+1-1+0 W+0-1+1 F+1-1+1"},
    {1, {-1, 1, 1}, false, "W+1+0+0 W+0+0-1 W+1+0-1 W+0-1+0 W+1-1+0 W+0-1-1 F+1-1-1"},
    {1, {-1, -1, 1}, false, "W+1+0+0 W+0+0-1 W+1+0-1 W+0+1+0 W+1+1+0 W+0+1-1 F+1+1-1"},
    {1, {1, 1, 1}, false, "W-1+0+0 W+0+0-1 W-1+0-1 W+0-1+0 W+0-1-1 W-1-1+0 F-1-1-1"},
    {1, {1, -1, 1}, false, "W-1+0+0 W+0+0-1 W-1+0-1 W+0+1+0 W+0+1-1 W-1+1+0 F-1+1-1"},
    {1, {1, -1, -1}, false, "W-1+0+0 W+0+0+1 W-1+0+1 W+0+1+0 W+0+1+1 W-1+1+0 F-1+1+1"},
    {1, {1, 1, -1}, false, "W-1+0+0 W+0+0+1 W-1+0+1 W+0-1+0 W+0-1+1 W-1-1+0 F-1-1+1"},
    {2, {1, 1, 1}, false, "F-1+0+0 F+0+0-1 W+0+0+1 W+1+0+0 F+0-1+0 W+0+1+0"},
    {2, {1, -1, 1}, false, "F-1+0+0 F+0+0-1 W+0+0+1 W+1+0+0 F+0+1+0 W+0-1+0"},
    {2, {1, -1, -1}, false, "F-1+0+0 F+0+0+1 W+0+0-1 W+1+0+0 F+0+1+0 W+0-1+0"},
    {2, {1, 1, -1}, false, "F-1+0+0 F+0+0+1 W+0+0-1 W+1+0+0 F+0-1+0 W+0+1+0"},
    {2, {-1, 1, 1}, false, "F+1+0+0 F+0+0-1 W-1+0+0 W+0+0+1 F+0-1+0 W+0+1+0"},
    {2, {-1, -1, 1}, false, "F+1+0+0 F+0+0-1 W-1+0+0 W+0+0+1 F+0+1+0 W+0-1+0"},
    {2, {-1, 1, -1}, false, "F+1+0+0 F+0+0+1 W-1+0+0 W+0+0-1 F+0-1+0 W+0+1+0"},
    {2, {-1, -1, -1}, false, "F+1+0+0 F+0+0+1 W-1+0+0 W+0+0-1 F+0+1+0 W+0-1+0"},
    {3, {0, 1, -1}, false, "W-1+0+0 W+1+0+0 W+0+0+1 N+1+0+1 N-1+0+1 F+0-1+1 W+0-1+0"},
    {3, {0, -1, -1}, false, "W-1+0+0 W+1+0+0 W+0+0+1 N+1+0+1 N-1+0+1 F+0+1+1 W+0+1+0"},
    {3, {0, -1, 1}, false, "W-1+0+0 W+1+0+0 W+0+0-1 N+1+0-1 N-1+0-1 F+0+1-1 N+0+1+0"},
    {3, {0, 1, 1}, false, "W-1+0+0 W+1+0+0 W+0+0-1 N+1+0-1 N-1+0-1 F+0-1-1 N+0-1+0"},
    {3, {-1, 0, -1}, false, "W+0+1+0 W+0-1+0 W+0+0+1 F+1+0+1 W+1+0+0 N-1+0+0 N+0+0-1"},
    {3, {1, 0, -1}, false, "W+0+1+0 W+0-1+0 W+0+0+1 F-1+0+1 W-1+0+0 N+1+0+0 N+0+0-1"},
    {3, {-1, 0, 1}, false, "W+0+1+0 W+0-1+0 W+0+0-1 F+1+0-1 W+1+0+0 N-1+0+0 N+0+0+1"},
    {3, {1, 0, 1}, false, "W+0+1+0 W+0-1+0 W+0+0-1 F-1+0-1 W-1+0+0 N+1+0+0 N+0+0+1"},
    {3, {-1, -1, 0}, false, "W+0+0+1 W+0+0-1 W+0+1+0 N+1+0+0 F+1+1+0"},
    {3, {1, -1, 0}, false, "W+0+0+1 W+0+0-1 W+0+1+0 N-1+0+0 F-1+1+0"},
    {3, {1, 1, 0}, false, "W+0+0+1 W+0+0-1 W+0-1+0 N-1+0+0 F-1-1+0"},
    {3, {-1, 1, 0}, false, "W+0+0+1 W+0+0-1 W+0-1+0 N+1+0+0 F+1-1+0"},
    {4, {0, -1, -1}, true, "W-1+0+0 W+1+0+0 W+0+0-1 F+0+0+1 F+0+1+0 W+0-1+0 W-1-1+0 W+1-1+0"},
    {4, {0, 1, -1}, true, "W-1+0+0 W+1+0+0 W+0+0-1 F+0+0+1 F+0-1+0 W+0+1+0 W-1+1+0 W+1+1+0"},
    {4, {0, -1, 1}, true, "W+1+0+0 W-1+0+0 W+0+0+1 F+0+0-1 F+0+1+0 W-1-1+0 W+0-1+0 W+1-1+0"},
    {4, {0, 1, 1}, true, "W+1+0+0 W-1+0+0 W+0+0+1 F+0+0-1 F+0-1+0 W-1+1+0 W+0+1+0 W+1+1+0"},
    {4, {1, 0, 1}, true, "F-1+0+0 W+1+0+0 W+0+1+0 W+0-1+0 F+0+0-1 W+0+0+1 W+0+1+1 W+0-1+1"},
    {4, {1, 0, -1}, true, "F-1+0+0 W+1+0+0 W+0+1+0 W+0-1+0 F+0+0+1 W+0+0-1 W+0+1-1 W+0-1-1"},
    {4, {-1, 0, 1}, true, "F+1+0+0 W-1+0+0 W+0-1+0 W+0+1+0 F+0+0-1 W+0+0+1 W+0+1+1 W+0-1+1"},
    {4, {-1, 0, -1}, true, "F+1+0+0 W-1+0+0 W+0-1+0 W+0+1+0 F+0+0+1 W+0+0-1 W+0+1-1 W+0-1-1"},
    {4, {1, 1, 0}, true, "F+0-1+0 W+0+0-1 W+0+0+1 W+0+1+0 F-1+0+0 F-1-1+0"},
    {4, {-1, 1, 0}, true, "F+0-1+0 W+0+0-1 W+0+0+1 W+0+1+0 F+1+0+0 F+1-1+0"},
    {4, {1, -1, 0}, true, "F+0+1+0 W+0+0-1 W+0+0+1 W+0-1+0 F-1+0+0 F-1+1+0"},
    {4, {-1, -1, 0}, true, "F+0+1+0 W+0+0-1 W+0+0+1 W+0-1+0 F+1+0+0 F+1+1+0"},
  };
  count = int(sizeof(rules) / sizeof(rules[0]));
  return rules;
}

// Material (dynamics id) of cell (x, y, z) of an nx*ny*nz map.  Periodic axes
// wrap.  Beyond a non-periodic face lies the first overlap layer: its
// materials come from `ring`, a padded (nx+2)*(ny+2)*(nz+2) map whose core
// entries are unused (the materials of the neighbouring blocks' cells, 0 where
// the domain ends), or are 0 when there is no ring.  Deeper than one layer is
// outside.
inline int material_at(const int n[3], const int periodic[3], const uint8_t* m, int x, int y, int z,
                       const uint8_t* ring = nullptr)
{
  int r[3] = {x, y, z};
  bool overlap = false;
  for (int d = 0; d < 3; ++d) {
    if (r[d] < 0 || r[d] >= n[d]) {
      if (!periodic[d]) {
        if (!ring || r[d] < -1 || r[d] > n[d]) return 0;
        overlap = true;
      } else {
        r[d] = (r[d] % n[d] + n[d]) % n[d];
      }
    }
  }
  if (overlap) {
    return ring[(int64_t(r[0] + 1) * (n[1] + 2) + (r[1] + 1)) * (n[2] + 2) + (r[2] + 1)];
  }
  return m[(int64_t(r[0]) * n[1] + r[1]) * n[2] + r[2]];
}

// Type and normal of wall cell (x, y, z); type kTypeFlat with a zero normal
// when no pattern matches.  `ring` as for material_at.
inline void classify_boundary_cell(const int n[3], const int periodic[3], const uint8_t* m, int x, int y, int z,
                                   int& type, int normal[3], const uint8_t* ring = nullptr)
{
  int count = 0;
  const NormalRule* rules = normal_rules(count);
  type = kTypeFlat;
  normal[0] = normal[1] = normal[2] = 0;
  bool found = false;
  for (int r = 0; r < count; ++r) {
    const NormalRule& rule = rules[r];
    if (found && !rule.replaces) continue;
    bool ok = true;
    for (const char* p = rule.pattern; *p && ok;) {
      const char pred = *p;
      int off[3];
      for (int d = 0; d < 3; ++d) {
        const int sign = p[1 + 2 * d] == '-' ? -1 : 1;
        off[d] = sign * (p[2 + 2 * d] - '0');
      }
      p += 7;
      while (*p == ' ') ++p;
      const int mat = material_at(n, periodic, m, x + off[0], y + off[1], z + off[2], ring);
      const bool fluid = mat == 1, outside = mat == 0;
      ok = pred == 'W' ? (!fluid && !outside) : (pred == 'F' ? fluid : !fluid);
    }
    if (!ok) continue;
    type = rule.type;
    for (int d = 0; d < 3; ++d) normal[d] = rule.normal[d];
    found = true;
  }
}

}  // namespace lbmh
