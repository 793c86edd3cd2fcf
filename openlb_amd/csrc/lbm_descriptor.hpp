// Discrete velocity sets D3Q19 and D3Q27 as compile-time tables.
//
// Order, weights and opposites are those of OpenLB's descriptors
// (src/descriptor/definition/common.h:120-166 for D3Q19, :313-366 for D3Q27):
// index 0 is the rest velocity, i and i + q/2 are opposite for i >= 1, and the
// first half holds the velocities whose leading non-zero component is -1.
// The tables must match the reference's because the population layout of
// every checkpoint and host upload is indexed by them.
#pragma once

#include <cstdint>

namespace lbmh {

template <int Q>
struct Lattice;

template <>
struct Lattice<19> {
  static constexpr int q = 19;
  static constexpr int half = 9;
  static constexpr int c[19][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
    {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
  // weights as (numerator, denominator): rest 1/3, axis 1/18, diagonal 1/36
  static constexpr int tnum[19] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int tden[19] = {3, 18, 18, 18, 36, 36, 36, 36, 36, 36,
                                   18, 18, 18, 36, 36, 36, 36, 36, 36};
};

template <>
struct Lattice<27> {
  static constexpr int q = 27;
  static constexpr int half = 13;
  static constexpr int c[27][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
    {-1, 0, 1}, {0, -1, -1}, {0, -1, 1}, {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
  // rest 8/27, axis 2/27, face diagonal 1/54, space diagonal 1/216
  static constexpr int tnum[27] = {8, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                   2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int tden[27] = {27, 27, 27, 27, 54, 54, 54, 54, 54, 54, 216, 216, 216, 216,
                                   27, 27, 27, 54, 54, 54, 54, 54, 54, 216, 216, 216, 216};
};

template <int Q>
constexpr int opposite(int i) { return i == 0 ? 0 : (i <= Lattice<Q>::half ? i + Lattice<Q>::half : i - Lattice<Q>::half); }

template <typename V, int Q>
__host__ __device__ constexpr V weight(int i) { return V(Lattice<Q>::tnum[i]) / V(Lattice<Q>::tden[i]); }

// Neighbour distance of velocity i in a padded block with strides (sx, sy, 1):
// BlockStructureD::getNeighborDistance (src/core/blockStructure.h:179-182).
template <int Q>
constexpr int64_t neighbor_distance(int i, int64_t sx, int64_t sy)
{
  return Lattice<Q>::c[i][0] * sx + Lattice<Q>::c[i][1] * sy + Lattice<Q>::c[i][2];
}

}  // namespace lbmh
