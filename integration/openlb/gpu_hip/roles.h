/* What a GPU_HIP block does with each OpenLB dynamics and post-processor type.
 *
 * The device lattice runs a fixed set of cell behaviours (include/lbmh.h):
 * one bulk collision per block (BGK, Guo-forced BGK, TRT, forced TRT,
 * Smagorinsky-BGK), NoDynamics, BounceBack, BounceBackVelocity walls, the
 * LocalVelocity family of boundaries, LocalPressure faces and the
 * InterpolatedVelocity / InterpolatedPressure faces with their plane
 * finite-difference post-processors.  Each OpenLB type is mapped here, from
 * its template arguments, to one of those roles; any other type is
 * Role::Unsupported and a block holding it refuses to simulate.
 */
#ifndef OLB_GPU_HIP_ROLES_H
#define OLB_GPU_HIP_ROLES_H

#include <array>
#include <type_traits>

namespace olb {

namespace gpu::hip {

enum class Role : std::uint8_t {
  Unsupported,
  None,             // NoDynamics                       src/dynamics/dynamics.h:50-62
  Bulk,             // one of the bulk collisions
  BounceBack,       // BounceBack                       dynamics.h:472-483
  MovingWall,       // BounceBackVelocity               dynamics.h:550-555
  VelocityFlat,     // LocalVelocity face               src/boundary/localVelocity3D.h:49-80
  VelocityFixed,    // LocalVelocity outer edge/corner  localVelocity3D.h:82-85, 92-95
  VelocityInnerCorner,  //                              localVelocity3D.h:87-90
  VelocityInnerEdge,    //                              localVelocity3D.h:97-100
  PressureFlat,         // LocalPressure face               src/boundary/localPressure3D.h:44-55
  InterpVelocityFlat,   // InterpolatedVelocity face        src/boundary/interpolatedVelocity3D.h:44-48
  InterpPressureFlat    // InterpolatedPressure face        src/boundary/interpolatedPressure3D.h:44-66
};

/// Boundary types in the numbering of lbmh_classify_boundary
enum BoundaryType : int { kFlat = 0, kExternalCorner = 1, kInternalCorner = 2, kExternalEdge = 3, kInternalEdge = 4 };

/// Normal of an edge given as (plane, normal1, normal2), the encoding of
/// boundaryhelper::promiseForNormalSpecial (src/boundary/setBoundary3D.h:586-626)
constexpr std::array<int,3> edgeNormal(int plane, int normal1, int normal2)
{
  return plane == 0 ? std::array<int,3>{0, normal1, normal2}
       : plane == 1 ? std::array<int,3>{normal2, 0, normal1}
       :              std::array<int,3>{normal1, normal2, 0};
}

struct RoleInfo {
  Role role = Role::Unsupported;
  int bulkKind = -1;                 // LBMH_BGK ... for Role::Bulk
  std::array<int,3> normal{0, 0, 0}; // boundary roles that carry their normal
};

template <typename T, typename DESCRIPTOR, typename DYNAMICS>
struct RoleOf {
  static constexpr RoleInfo info() { return {}; }
};

/// Dynamics OpenLB runs through a generated common-subexpression kernel come
/// wrapped in CSE<...> (src/cse/wrapper.h:40-57); their role is the wrapped one
template <typename DYNAMICS>
struct Unwrapped {
  using type = DYNAMICS;
};

template <typename DYNAMICS>
struct Unwrapped<CSE<DYNAMICS>> {
  using type = DYNAMICS;
};

template <typename T, typename DESCRIPTOR, typename DYNAMICS>
constexpr RoleInfo roleOf()
{
  using D = typename Unwrapped<DYNAMICS>::type;
  if constexpr (std::is_same_v<D, NoDynamics<T,DESCRIPTOR>>) {
    return {Role::None};
  } else if constexpr (std::is_same_v<D, BGKdynamics<T,DESCRIPTOR>>) {
    return {Role::Bulk, LBMH_BGK};
  } else if constexpr (std::is_same_v<D, ForcedBGKdynamics<T,DESCRIPTOR>>) {
    return {Role::Bulk, LBMH_FORCED_BGK};
  } else if constexpr (std::is_same_v<D, TRTdynamics<T,DESCRIPTOR>>) {
    return {Role::Bulk, LBMH_TRT};
  } else if constexpr (std::is_same_v<D, ForcedTRTdynamics<T,DESCRIPTOR>>) {
    return {Role::Bulk, LBMH_FORCED_TRT};
  } else if constexpr (std::is_same_v<D, SmagorinskyBGKdynamics<T,DESCRIPTOR>>) {
    return {Role::Bulk, LBMH_SMAGORINSKY};
  } else if constexpr (std::is_same_v<D, BounceBack<T,DESCRIPTOR>>) {
    return {Role::BounceBack};
  } else if constexpr (std::is_same_v<D, BounceBackVelocity<T,DESCRIPTOR>>) {
    return {Role::MovingWall};
  } else if constexpr (std::is_same_v<D, BGKdynamics<T,DESCRIPTOR,momenta::FixedVelocityBoundaryTuple>>) {
    return {Role::VelocityFixed};
  } else {
    return RoleOf<T,DESCRIPTOR,D>::info();
  }
}

/// LocalVelocity faces: CombinedRLBdynamics with the regularized velocity momenta
template <typename T, typename DESCRIPTOR, int direction, int orientation>
struct RoleOf<T,DESCRIPTOR,CombinedRLBdynamics<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>,
                                               momenta::RegularizedVelocityBoundaryTuple<direction,orientation>>> {
  static constexpr RoleInfo info()
  {
    RoleInfo r{Role::VelocityFlat};
    r.normal[direction] = orientation;
    return r;
  }
};

template <typename T, typename DESCRIPTOR, int x, int y, int z>
struct RoleOf<T,DESCRIPTOR,CombinedRLBdynamics<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>,
                                               momenta::InnerCornerVelocityTuple3D<x,y,z>>> {
  static constexpr RoleInfo info() { return {Role::VelocityInnerCorner, -1, {x, y, z}}; }
};

template <typename T, typename DESCRIPTOR, int plane, int normal1, int normal2>
struct RoleOf<T,DESCRIPTOR,CombinedRLBdynamics<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>,
                                               momenta::InnerEdgeVelocityTuple3D<plane,normal1,normal2>>> {
  static constexpr RoleInfo info() { return {Role::VelocityInnerEdge, -1, edgeNormal(plane, normal1, normal2)}; }
};

/// LocalPressure faces: CombinedRLBdynamics with the regularized pressure momenta
template <typename T, typename DESCRIPTOR, int direction, int orientation>
struct RoleOf<T,DESCRIPTOR,CombinedRLBdynamics<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR>,
                                               momenta::RegularizedPressureBoundaryTuple<direction,orientation>>> {
  static constexpr RoleInfo info()
  {
    RoleInfo r{Role::PressureFlat};
    r.normal[direction] = orientation;
    return r;
  }
};

/// Interpolated faces: BGK with the Dirichlet velocity / pressure momenta
/// (the mixin's exchange_momenta), completed by a PlaneFdBoundaryProcessor3D
template <typename T, typename DESCRIPTOR, int direction, int orientation>
struct RoleOf<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR,momenta::BasicDirichletVelocityBoundaryTuple<direction,orientation>>> {
  static constexpr RoleInfo info()
  {
    RoleInfo r{Role::InterpVelocityFlat};
    r.normal[direction] = orientation;
    return r;
  }
};

template <typename T, typename DESCRIPTOR, int direction, int orientation>
struct RoleOf<T,DESCRIPTOR,BGKdynamics<T,DESCRIPTOR,momenta::BasicDirichletPressureBoundaryTuple<direction,orientation>>> {
  static constexpr RoleInfo info()
  {
    RoleInfo r{Role::InterpPressureFlat};
    r.normal[direction] = orientation;
    return r;
  }
};

/// Post-processors the device step runs itself
struct PostInfo {
  bool handled = false;
  int type = -1;                     // kExternalEdge / kExternalCorner
  std::array<int,3> normal{0, 0, 0};
};

template <typename OPERATOR>
struct PostOf {
  static constexpr PostInfo info() { return {}; }
};

/// Post-processors with a generated kernel come wrapped in CSE_O<...>
/// (src/cse/wrapper.h:165-185), like dynamics in CSE<...>
template <typename OPERATOR, typename DESCRIPTOR>
struct PostOf<CSE_O<OPERATOR,DESCRIPTOR>> : PostOf<OPERATOR> { };

template <typename T, typename DESCRIPTOR, int plane, int normal1, int normal2>
struct PostOf<OuterVelocityEdgeProcessor3D<T,DESCRIPTOR,plane,normal1,normal2>> {
  static constexpr PostInfo info() { return {true, kExternalEdge, edgeNormal(plane, normal1, normal2)}; }
};

template <typename T, typename DESCRIPTOR, int x, int y, int z>
struct PostOf<OuterVelocityCornerProcessor3D<T,DESCRIPTOR,x,y,z>> {
  static constexpr PostInfo info() { return {true, kExternalCorner, {x, y, z}}; }
};

/// The plane finite-difference processor of the interpolated faces
/// (src/boundary/boundaryPostProcessors3D.h:41-70)
template <typename T, typename DESCRIPTOR, int direction, int orientation>
struct PostOf<PlaneFdBoundaryProcessor3D<T,DESCRIPTOR,direction,orientation>> {
  static constexpr PostInfo info()
  {
    PostInfo p{true, kFlat};
    p.normal[direction] = orientation;
    return p;
  }
};

}

}

#endif
