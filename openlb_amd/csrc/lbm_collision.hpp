// Fused per-cell collisions on shifted populations fb_i = f_i - t_i.
//
// The arithmetic is organised around opposite pairs (i, i + q/2): for a pair
// c.u flips sign, so the equilibrium splits into an even part
//   t_i (drho - 1.5 rho u^2 + 4.5 rho (c_i.u)^2)
// and an odd part t_i 3 rho (c_i.u), both computed from the deviation
// drho = rho - 1 = sum fb_i.  This keeps the FP32/FP16-stored values exact to
// the last stored bit (no 1 + small cancellation) and halves the equilibrium
// work.  Each function is the physics of one OpenLB operator:
//   bgk          collision::BGK + equilibria::SecondOrder   src/dynamics/collision.h:57-80, lbm.h:50-64
//   trt          collision::TRT                             src/dynamics/collision.h:209-247
//   guo          forcing::Guo / lbm::addExternalForce       src/dynamics/forcing.h:37-74, lbm.h:545-566
//   smagorinsky  SmagorinskyEffectiveOmega<BGK>             src/dynamics/collisionLES.h:58-89
//   revert       collision::Revert (BounceBack)             src/dynamics/collision.h:293-312
//   moving_wall  NguyenLaddCorrection<Revert>               src/dynamics/collision.h:323-348
#pragma once

#include "lbm_descriptor.hpp"

namespace lbmh {

// Kinds of the bulk collision (cells with dynamics id 1).
enum BulkKind : int { kBGK = 0, kForcedBGK = 1, kTRT = 2, kForcedTRT = 3, kSmagorinsky = 4 };

template <typename V>
struct BulkParams {
  V omega;      // 1 / tau
  V omega2;     // TRT odd-moment rate 1/(magic/(tau - 1/2) + 1/2)
  V guo;        // 1 - omega/2
  V force[3];   // uniform body force
  V half_force[3];
  // Smagorinsky: omega_eff = 1 / (sm_a sqrt(sm_b |Pi_neq| / rho + sm_k tau^2) + tau/2),
  // sm_b already multiplied by C^2
  V sm_a, sm_b, sm_k, tau, half_tau;
};

template <int Q, typename V>
__device__ __forceinline__ void moments(const V (&f)[Q], V& drho, V (&j)[3])
{
  using L = Lattice<Q>;
  V s = f[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) s += f[i];
  drho = s;
  j[0] = j[1] = j[2] = V(0);
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    const V d = f[i] - f[i + L::half];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (L::c[i][a] > 0) j[a] += d;
      if (L::c[i][a] < 0) j[a] -= d;
    }
  }
}

template <int Q, typename V>
__device__ __forceinline__ V dot_c(int i, const V (&u)[3])
{
  using L = Lattice<Q>;
  V s = V(0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (L::c[i][a] > 0) s += u[a];
    if (L::c[i][a] < 0) s -= u[a];
  }
  return s;
}

// Equilibrium halves for pair (i, opp(i)): feq_i = even + odd, feq_opp = even - odd.
template <int Q, typename V>
__device__ __forceinline__ void equilibrium_pair(int i, V A, V rho, const V (&u)[3], V& even, V& odd)
{
  const V t = weight<V, Q>(i);
  const V cu = dot_c<Q, V>(i, u);
  even = t * (A + V(4.5) * rho * cu * cu);
  odd = t * (V(3) * rho * cu);
}

// BGK relaxation towards the second-order equilibrium at (rho, u).
template <int Q, typename V>
__device__ __forceinline__ void bgk(V (&f)[Q], V drho, V rho, const V (&u)[3], V omega)
{
  using L = Lattice<Q>;
  const V usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const V A = drho - V(1.5) * rho * usq;
  f[0] += omega * (weight<V, Q>(0) * A - f[0]);
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    V even, odd;
    equilibrium_pair<Q, V>(i, A, rho, u, even, odd);
    f[i] += omega * (even + odd - f[i]);
    f[i + L::half] += omega * (even - odd - f[i + L::half]);
  }
}

// Two-relaxation-time collision: omega on the even, omega2 on the odd part.
template <int Q, typename V>
__device__ __forceinline__ void trt(V (&f)[Q], V drho, V rho, const V (&u)[3], V omega, V omega2)
{
  using L = Lattice<Q>;
  const V usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const V A = drho - V(1.5) * rho * usq;
  f[0] -= omega * (f[0] - weight<V, Q>(0) * A);
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    V even, odd;
    equilibrium_pair<Q, V>(i, A, rho, u, even, odd);
    const V fp = V(0.5) * (f[i] + f[i + L::half]);
    const V fm = V(0.5) * (f[i] - f[i + L::half]);
    const V de = omega * (fp - even);
    const V dodd = omega2 * (fm - odd);
    f[i] -= de + dodd;
    f[i + L::half] -= de - dodd;
  }
}

// Guo source term, added after relaxation with the shifted velocity u.
template <int Q, typename V>
__device__ __forceinline__ void guo(V (&f)[Q], V rho, const V (&u)[3], const BulkParams<V>& p)
{
  using L = Lattice<Q>;
  const V uF = u[0] * p.force[0] + u[1] * p.force[1] + u[2] * p.force[2];
  const V s = p.guo * rho;
  f[0] += s * weight<V, Q>(0) * (V(-3) * uF);
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    const V t = weight<V, Q>(i) * s;
    const V cu = dot_c<Q, V>(i, u);
    const V cF = dot_c<Q, V>(i, p.force);
    const V even = t * (V(9) * cu * cF - V(3) * uF);
    const V odd = t * (V(3) * cF);
    f[i] += even + odd;
    f[i + L::half] += even - odd;
  }
}

// Effective relaxation rate of the Smagorinsky model from the non-equilibrium stress.
template <int Q, typename V>
__device__ __forceinline__ V smagorinsky_omega(const V (&f)[Q], V drho, V rho, const V (&u)[3],
                                               const BulkParams<V>& p)
{
  using L = Lattice<Q>;
  // sum_i c_ia c_ib fb_i over pairs (even in c)
  V P[6] = {V(0), V(0), V(0), V(0), V(0), V(0)};  // xx xy xz yy yz zz
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    const V s = f[i] + f[i + L::half];
    const int cx = L::c[i][0], cy = L::c[i][1], cz = L::c[i][2];
    if (cx) P[0] += s;
    if (cy) P[3] += s;
    if (cz) P[5] += s;
    if (cx * cy > 0) P[1] += s;
    if (cx * cy < 0) P[1] -= s;
    if (cx * cz > 0) P[2] += s;
    if (cx * cz < 0) P[2] -= s;
    if (cy * cz > 0) P[4] += s;
    if (cy * cz < 0) P[4] -= s;
  }
  const V d3 = drho * V(1.0 / 3.0);
  const V xx = P[0] - rho * u[0] * u[0] - d3;
  const V yy = P[3] - rho * u[1] * u[1] - d3;
  const V zz = P[5] - rho * u[2] * u[2] - d3;
  const V xy = P[1] - rho * u[0] * u[1];
  const V xz = P[2] - rho * u[0] * u[2];
  const V yz = P[4] - rho * u[1] * u[2];
  const V nsq = xx * xx + yy * yy + zz * zz + V(2) * (xy * xy + xz * xz + yz * yz);
  const V norm = sqrt(nsq);
  const V inner = p.sm_b * norm / rho + p.sm_k * (p.tau * p.tau);
  return V(1) / (p.sm_a * sqrt(inner) + p.half_tau);
}

template <int Q, typename V>
__device__ __forceinline__ void revert(V (&f)[Q])
{
  using L = Lattice<Q>;
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    const V tmp = f[i];
    f[i] = f[i + L::half];
    f[i + L::half] = tmp;
  }
}

// Revert, then the Nguyen-Ladd momentum correction 2 rho_w t_i (c_i.u_w) / cs^2.
template <int Q, typename V>
__device__ __forceinline__ void moving_wall(V (&f)[Q], V rho_w, const V (&uw)[3])
{
  using L = Lattice<Q>;
  revert<Q, V>(f);
#pragma unroll
  for (int i = 1; i <= L::half; ++i) {
    const V corr = V(6) * rho_w * weight<V, Q>(i) * dot_c<Q, V>(i, uw);
    f[i] += corr;
    f[i + L::half] -= corr;
  }
}

// Bulk collision of kind KIND on one cell.
template <int Q, typename V, int KIND>
__device__ __forceinline__ void collide_bulk(V (&f)[Q], const BulkParams<V>& p)
{
  V drho, j[3];
  moments<Q, V>(f, drho, j);
  const V rho = V(1) + drho;
  const V inv = V(1) / rho;
  V u[3] = {j[0] * inv, j[1] * inv, j[2] * inv};
  if constexpr (KIND == kForcedBGK || KIND == kForcedTRT) {
    u[0] += p.half_force[0];
    u[1] += p.half_force[1];
    u[2] += p.half_force[2];
  }
  if constexpr (KIND == kBGK || KIND == kForcedBGK) {
    bgk<Q, V>(f, drho, rho, u, p.omega);
  } else if constexpr (KIND == kTRT || KIND == kForcedTRT) {
    trt<Q, V>(f, drho, rho, u, p.omega, p.omega2);
  } else {
    const V w = smagorinsky_omega<Q, V>(f, drho, rho, u, p);
    bgk<Q, V>(f, drho, rho, u, w);
  }
  if constexpr (KIND == kForcedBGK || KIND == kForcedTRT) {
    guo<Q, V>(f, rho, u, p);
  }
}

}  // namespace lbmh
