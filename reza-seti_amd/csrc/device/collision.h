// Device-side lattice-Boltzmann arithmetic shared by the kernel families:
// storage traits, the reference FP64 collision, the generic (shifted / forced)
// collision, equilibria and the reference Zou-He rule.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../lattice.h"

namespace lbm {
namespace dev {

constexpr int WALL_REF = 0, WALL_BB = 1;
constexpr int INLET_ZOUHE = 0, INLET_NEE = 1;
constexpr int OUTLET_REF = 0, OUTLET_NEE = 1;

// ---------------------------------------------------------------- storage
// double: f unshifted. float / __half: f - w_q, arithmetic in float.
template <class S>
struct Pop;
template <>
struct Pop<double> {
  using C = double;
  static constexpr bool shifted = false;
  __device__ static C ld(const void* base, int64_t i) { return static_cast<const double*>(base)[i]; }
  __device__ static void st(void* base, int64_t i, C v) { static_cast<double*>(base)[i] = v; }
  // Nontemporal (cache-bypassing) variants for streamed data.
  __device__ static C ld_nt(const void* base, int64_t i) {
    return __builtin_nontemporal_load(static_cast<const double*>(base) + i);
  }
  __device__ static void st_nt(void* base, int64_t i, C v) {
    __builtin_nontemporal_store(v, static_cast<double*>(base) + i);
  }
};
template <>
struct Pop<float> {
  using C = float;
  static constexpr bool shifted = true;
  __device__ static C ld(const void* base, int64_t i) { return static_cast<const float*>(base)[i]; }
  __device__ static void st(void* base, int64_t i, C v) { static_cast<float*>(base)[i] = v; }
  __device__ static C ld_nt(const void* base, int64_t i) {
    return __builtin_nontemporal_load(static_cast<const float*>(base) + i);
  }
  __device__ static void st_nt(void* base, int64_t i, C v) {
    __builtin_nontemporal_store(v, static_cast<float*>(base) + i);
  }
};
template <>
struct Pop<__half> {
  using C = float;
  static constexpr bool shifted = true;
  __device__ static C ld(const void* base, int64_t i) {
    return __half2float(static_cast<const __half*>(base)[i]);
  }
  __device__ static void st(void* base, int64_t i, C v) { static_cast<__half*>(base)[i] = __float2half_rn(v); }
  __device__ static C ld_nt(const void* base, int64_t i) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(static_cast<const unsigned short*>(base) + i)));
  }
  __device__ static void st_nt(void* base, int64_t i, C v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)), static_cast<unsigned short*>(base) + i);
  }
};

// Element at byte offset `off` of a wave-uniform base: the scalar base +
// 32-bit vector offset addressing form, no 64-bit address arithmetic per
// lane (buffers of fewer than 2^32 bytes).
template <class S, bool NT>
__device__ inline typename Pop<S>::C ld_off(const void* base, uint32_t off) {
  const char* b = static_cast<const char*>(base) + off;
  if constexpr (std::is_same<S, __half>::value) {
    const unsigned short* q = reinterpret_cast<const unsigned short*>(b);
    return __half2float(__ushort_as_half(NT ? __builtin_nontemporal_load(q) : *q));
  } else {
    const S* q = reinterpret_cast<const S*>(b);
    if constexpr (NT) return __builtin_nontemporal_load(q);
    else return *q;
  }
}
template <class S, bool NT>
__device__ inline void st_off(void* base, uint32_t off, typename Pop<S>::C v) {
  char* b = static_cast<char*>(base) + off;
  if constexpr (std::is_same<S, __half>::value) {
    unsigned short* q = reinterpret_cast<unsigned short*>(b);
    const unsigned short h = __half_as_ushort(__float2half_rn(v));
    if constexpr (NT) __builtin_nontemporal_store(h, q);
    else *q = h;
  } else {
    S* q = reinterpret_cast<S*>(b);
    if constexpr (NT) __builtin_nontemporal_store(S(v), q);
    else *q = S(v);
  }
}

// ------------------------------------------------- reference FP64 collision
// Velocity sums grouped exactly as Simulation.cpp:100-134 (D3Q19) and
// :149-207 (D3Q27); each line is a left-to-right sum divided by rho.
template <class L>
struct RefVelocity;
template <>
struct RefVelocity<D3Q19> {
  __device__ static void get(const double* f, double rho, double& ux, double& uy, double& uz) {
    ux = ((f[0] + f[6] + f[7] + f[8] + f[9]) - (f[1] + f[10] + f[11] + f[12] + f[13])) / rho;
    uy = ((f[2] + f[6] + f[10] + f[14] + f[15]) - (f[3] + f[7] + f[11] + f[16] + f[17])) / rho;
    uz = ((f[4] + f[8] + f[12] + f[14] + f[16]) - (f[5] + f[9] + f[13] + f[15] + f[17])) / rho;
  }
};
template <>
struct RefVelocity<D3Q27> {
  __device__ static void get(const double* f, double rho, double& ux, double& uy, double& uz) {
    ux = ((f[1] + f[7] + f[8] + f[11] + f[13] + f[19] + f[21] + f[23] + f[25]) -
          (f[2] + f[9] + f[10] + f[12] + f[14] + f[20] + f[22] + f[24] + f[26])) / rho;
    uy = ((f[3] + f[7] + f[9] + f[15] + f[16] + f[19] + f[21] + f[24] + f[26]) -
          (f[4] + f[8] + f[10] + f[17] + f[18] + f[20] + f[22] + f[23] + f[25])) / rho;
    uz = ((f[5] + f[11] + f[12] + f[15] + f[17] + f[19] + f[22] + f[23] + f[26]) -
          (f[6] + f[13] + f[14] + f[16] + f[18] + f[20] + f[21] + f[24] + f[25])) / rho;
  }
};

// m_eqdistributions[q] of Simulation.cpp:214-224.
template <class L>
__device__ inline double feq_ref(int q, double rho, double ux, double uy, double uz) {
  const double du = ux * L::cx[q] + uy * L::cy[q] + uz * L::cz[q];
  const double u_sqr = ux * ux + uy * uy + uz * uz;
  return L::w[q] * rho * (1.0 + 3.0 * du + 4.5 * du * du - 1.5 * u_sqr);
}

// Simulation::Collision for one node (Simulation.cpp:79-231).
template <class L>
__device__ inline void collide_ref(double* f, bool wall, double omega, double& rho, double& ux, double& uy,
                                   double& uz) {
  rho = 0.0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) rho += f[q];
  if (wall) {
    ux = 0.0;
    uy = 0.0;
    uz = 0.0;
  } else {
    RefVelocity<L>::get(f, rho, ux, uy, uz);
  }
  double feq[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) feq[q] = feq_ref<L>(q, rho, ux, uy, uz);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) f[q] = f[q] + omega * (feq[q] - f[q]);
}

// ----------------------------------------- generic collision (extensions)
// BGK with optional Guo body force and Brinkman damping alpha:
//   u = (j + F/2) / (rho (1 + alpha/2)),  F_tot = F - alpha rho u,
//   f* = f + omega (feq - f) + (1 - omega/2) w [3 (c - u).F_tot + 9 (c.u)(c.F_tot)].
// SHIFT: populations are f - w, and drho = sum(f - w) = rho - 1.
// c_q . (x, y, z) over the nonzero components of c_q only (q is a
// compile-time constant once the loops are unrolled). Written as
// x * cx + y * cy + z * cz, the zero terms would stay in the code: 0 * x is
// not foldable when x may be infinite or NaN.
template <class L, class C>
__device__ inline C cdot(int q, C x, C y, C z) {
  C r = C(0);
  bool any = false;
  auto add = [&](int c, C v) {
    if (c == 0) return;
    const C t = c > 0 ? v : -v;
    r = any ? r + t : t;
    any = true;
  };
  add(L::cx[q], x);
  add(L::cy[q], y);
  add(L::cz[q], z);
  return r;
}

// w_q (drho + rho P_q) (SHIFT) or w_q rho (1 + P_q), with
// P_q = 3 cu + 4.5 cu^2 - 1.5 u^2 evaluated as cu (3 + 4.5 cu) - 1.5 u^2.
template <class L, class C, bool SHIFT>
__device__ inline C geq_of(int q, C drho, C rho, C ux, C uy, C uz, C usq) {
#pragma clang fp contract(fast)
  const C cu = cdot<L, C>(q, ux, uy, uz);
  const C wq = C(L::w[q]);
  const C poly = cu * (C(3) + C(4.5) * cu) - C(1.5) * usq;
  if (SHIFT) return wq * (drho + rho * poly);
  return wq * rho * (C(1) + poly);
}

// A sum of terms selected at compile time that starts from its first term
// (0 + x is not foldable: it differs from x for x = -0).
template <class C>
struct Sum {
  C v = C(0);
  bool any = false;
  __device__ void add(C x) {
    v = any ? v + x : x;
    any = true;
  }
  __device__ void sub(C x) {
    v = any ? v - x : -x;
    any = true;
  }
  // += sign * x for a compile-time sign in {-1, 0, 1}
  __device__ void acc(int sign, C x) {
    if (sign > 0) add(x);
    if (sign < 0) sub(x);
  }
};

template <class L, class C>
__device__ inline void moments(const C* f, C& s, C& jx, C& jy, C& jz) {
#pragma clang fp contract(fast)
  Sum<C> S, X, Y, Z;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    S.add(f[q]);
    X.acc(L::cx[q], f[q]);
    Y.acc(L::cy[q], f[q]);
    Z.acc(L::cz[q], f[q]);
  }
  s = S.v;
  jx = X.v;
  jy = Y.v;
  jz = Z.v;
}

template <class L, class C, bool SHIFT, bool FORCE>
__device__ inline void collide_ext(C* f, C omega, const C* force, C alpha, bool zero_u, C& drho, C& ux, C& uy,
                                   C& uz) {
#pragma clang fp contract(fast)
  C s, jx, jy, jz;
  moments<L, C>(f, s, jx, jy, jz);
  const C rho = SHIFT ? C(1) + s : s;
  drho = s;
  C Fx = C(0), Fy = C(0), Fz = C(0);
  if (FORCE) {
    const C inv = C(1) / (rho * (C(1) + C(0.5) * alpha));
    ux = (jx + C(0.5) * force[0]) * inv;
    uy = (jy + C(0.5) * force[1]) * inv;
    uz = (jz + C(0.5) * force[2]) * inv;
    if (zero_u) ux = uy = uz = C(0);
    Fx = force[0] - alpha * rho * ux;
    Fy = force[1] - alpha * rho * uy;
    Fz = force[2] - alpha * rho * uz;
  } else {
    const C inv = C(1) / rho;
    ux = jx * inv;
    uy = jy * inv;
    uz = jz * inv;
    if (zero_u) ux = uy = uz = C(0);
  }
  const C usq = ux * ux + uy * uy + uz * uz;
  const C kf = C(1) - C(0.5) * omega;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const C e = geq_of<L, C, SHIFT>(q, s, rho, ux, uy, uz, usq);
    C out = f[q] + omega * (e - f[q]);
    if (FORCE) {
      const C cu = cdot<L, C>(q, ux, uy, uz);
      const C cF = cdot<L, C>(q, Fx, Fy, Fz);
      const C uF = ux * Fx + uy * Fy + uz * Fz;
      out += kf * C(L::w[q]) * (C(3) * (cF - uF) + C(9) * cu * cF);
    }
    f[q] = out;
  }
}

template <class L, class C, bool SHIFT>
__device__ inline void equilibrium(C* e, C rho, C ux, C uy, C uz) {
  const C usq = ux * ux + uy * uy + uz * uz;
  const C drho = SHIFT ? rho - C(1) : C(0);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) e[q] = geq_of<L, C, SHIFT>(q, drho, rho, ux, uy, uz, usq);
}

template <class L, class C, bool SHIFT>
__device__ inline void density_velocity(const C* f, C& rho, C& ux, C& uy, C& uz) {
  C s, jx, jy, jz;
  moments<L, C>(f, s, jx, jy, jz);
  rho = SHIFT ? C(1) + s : s;
  ux = jx / rho;
  uy = jy / rho;
  uz = jz / rho;
}

// Non-equilibrium extrapolation from the interior neighbour's post-stream
// populations G: f = feq(rho_b, u_b) + (G - feq(rho_G, u_G)). rho_b <= 0 takes
// rho_G; use_G_velocity takes u_b = u_G (pressure outlet).
template <class L, class C, bool SHIFT>
__device__ inline void nee(C* F, const C* G, C rho_b, C ux, C uy, C uz, bool use_G_velocity) {
  C rg, gx, gy, gz;
  density_velocity<L, C, SHIFT>(G, rg, gx, gy, gz);
  if (use_G_velocity) {
    ux = gx;
    uy = gy;
    uz = gz;
  }
  if (rho_b <= C(0)) rho_b = rg;
  C eb[L::Q], eg[L::Q];
  equilibrium<L, C, SHIFT>(eb, rho_b, ux, uy, uz);
  equilibrium<L, C, SHIFT>(eg, rg, gx, gy, gz);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) F[q] = eb[q] + (G[q] - eg[q]);
}

// ------------------------------------------------------------- boundaries
// Boundary::ApplyBc (Boundary.cpp:111-289): six independent branches on the
// sign of each normal component, in source order, with the source's
// expression order. Branches for +-y hard-code the velocity 1.0, as in the
// reference. SHIFT adds the sum of the weights involved (exactly 1 for every
// branch) to the density numerator.
template <class C, bool SHIFT>
__device__ inline void zouhe_ref(C* m, C vel, const double* normal) {
  C v0, v1, v2;
  if (normal[0] > 0) {
    v0 = vel; v1 = C(0.0); v2 = C(0.0);
    const C S = m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                C(2) * (m[0] + m[6] + m[7] + m[8] + m[9]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) + v0);
    C Nxy = (C(1.0) / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - (C(1.0) / 3) * (rhoN * v1);
    C Nxz = (C(1.0) / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - (C(1.0) / 3) * (rhoN * v2);
    m[1] = m[0] - (1 / C(3.0)) * rhoN * v0;
    m[10] = m[7] + (1 / C(6.0)) * rhoN * (-v0 + v1) - Nxy;
    m[11] = m[6] + (1 / C(6.0)) * rhoN * (-v0 - v1) + Nxy;
    m[13] = m[8] + (1 / C(6.0)) * rhoN * (-v0 - v2) + Nxz;
    m[12] = m[9] + (1 / C(6.0)) * rhoN * (-v0 + v2) - Nxz;
  }
  if (normal[0] < 0) {
    v0 = vel; v1 = C(0.0); v2 = C(0.0);
    const C S = m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                C(2) * (m[1] + m[10] + m[11] + m[12] + m[13]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) - v0);
    C Nxy = (C(1.0) / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - (C(1.0) / 3) * (rhoN * v1);
    C Nxz = (C(1.0) / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - (C(1.0) / 3) * (rhoN * v2);
    m[0] = m[1] + (1 / C(3.0)) * rhoN * v0;
    m[7] = m[10] + (1 / C(6.0)) * rhoN * (v0 - v1) + Nxy;
    m[6] = m[11] + (1 / C(6.0)) * rhoN * (v0 + v1) - Nxy;
    m[8] = m[13] + (1 / C(6.0)) * rhoN * (v0 + v2) - Nxz;
    m[9] = m[12] + (1 / C(6.0)) * rhoN * (v0 - v2) + Nxz;
  }
  if (normal[1] > 0) {
    v0 = C(0.0); v1 = C(1.0); v2 = C(0.0);
    const C S = m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                C(2) * (m[2] + m[6] + m[10] + m[14] + m[15]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) + v1);
    C Nyx = (C(1.0) / 2) * (m[0] + m[8] + m[9] - m[1] - m[12] - m[13]) - (C(1.0) / 3) * (rhoN * v0);
    C Nyz = (C(1.0) / 2) * (m[4] + m[8] + m[12] - m[5] - m[9] - m[13]) - (C(1.0) / 3) * (rhoN * v2);
    m[3] = m[2] - (1 / C(3.0)) * rhoN * v1;
    m[11] = m[6] + (1 / C(6.0)) * rhoN * (-v0 - v1) + Nyx;
    m[7] = m[10] + (1 / C(6.0)) * rhoN * (-v1 + v0) - Nyx;
    m[17] = m[14] + (1 / C(6.0)) * rhoN * (-v1 - v2) + Nyz;
    m[16] = m[15] + (1 / C(6.0)) * rhoN * (-v1 + v2) - Nyz;
  }
  if (normal[1] < 0) {
    v0 = C(0.0); v1 = C(1.0); v2 = C(0.0);
    const C S = m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                C(2) * (m[3] + m[7] + m[11] + m[16] + m[17]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) - v1);
    C Nyx = (C(1.0) / 2) * (m[0] + m[8] + m[9] - m[1] - m[12] - m[13]) - (C(1.0) / 3) * (rhoN * v0);
    C Nyz = (C(1.0) / 2) * (m[4] + m[8] + m[12] - m[5] - m[9] - m[13]) - (C(1.0) / 3) * (rhoN * v2);
    m[2] = m[3] + (1 / C(3.0)) * rhoN * v1;
    m[6] = m[11] + (1 / C(6.0)) * rhoN * (v0 + v1) - Nyx;
    m[10] = m[7] + (1 / C(6.0)) * rhoN * (v1 - v0) + Nyx;
    m[14] = m[17] + (1 / C(6.0)) * rhoN * (v1 + v2) - Nyz;
    m[15] = m[16] + (1 / C(6.0)) * rhoN * (v1 - v2) + Nyz;
  }
  if (normal[2] > 0) {
    v0 = C(0.0); v1 = C(0.0); v2 = vel;
    const C S = m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                C(2) * (m[3] + m[7] + m[11] + m[16] + m[17]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) + v2);
    C Nzx = (C(1.0) / 2) * (m[0] + m[6] + m[7] - m[1] - m[10] - m[11]) - (C(1.0) / 3) * (rhoN * v0);
    C Nzy = (C(1.0) / 2) * (m[2] + m[6] + m[10] - m[3] - m[7] - m[11]) - (C(1.0) / 3) * (rhoN * v1);
    m[5] = m[4] - (1 / C(3.0)) * rhoN * v2;
    m[9] = m[12] + (1 / C(6.0)) * rhoN * (v0 - v2) - Nzx;
    m[13] = m[8] + (1 / C(6.0)) * rhoN * (-v0 - v2) + Nzx;
    m[15] = m[16] + (1 / C(6.0)) * rhoN * (v1 - v2) - Nzy;
    m[17] = m[14] + (1 / C(6.0)) * rhoN * (-v1 - v2) + Nzy;
  }
  if (normal[2] < 0) {
    v0 = C(0.0); v1 = C(0.0); v2 = vel;
    const C S = m[0] + m[1] + m[2] + m[3] + m[6] + m[7] + m[10] + m[11] + m[18] +
                C(2) * (m[5] + m[9] + m[13] + m[15] + m[17]);
    const C rhoN = (SHIFT ? C(1) + S : S) / (C(1) - v2);
    C Nzx = (C(1.0) / 2) * (m[0] + m[6] + m[7] - m[1] - m[10] - m[11]) - (C(1.0) / 3) * (rhoN * v0);
    C Nzy = (C(1.0) / 2) * (m[2] + m[6] + m[10] - m[3] - m[7] - m[11]) - (C(1.0) / 3) * (rhoN * v1);
    m[4] = m[5] + (1 / C(3.0)) * rhoN * v2;
    m[8] = m[13] + (1 / C(6.0)) * rhoN * (v0 + v2) - Nzx;
    m[12] = m[9] + (1 / C(6.0)) * rhoN * (-v0 + v2) + Nzx;
    m[14] = m[17] + (1 / C(6.0)) * rhoN * (v1 + v2) - Nzy;
    m[16] = m[15] + (1 / C(6.0)) * rhoN * (-v1 + v2) + Nzy;
  }
}

}  // namespace dev
}  // namespace lbm
