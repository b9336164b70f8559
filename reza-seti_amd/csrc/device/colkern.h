// Per-node building blocks of the column-segment step kernel (pull with
// half-way bounce-back, non-equilibrium extrapolation, wall shear stress,
// collision), shared by the forward kernels (columns.hip) and their adjoint
// (adjoint.hip). Layout: columns.h.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../lattice.h"
#include "collision.h"
#include "columns.h"
#include "kernels.h"

namespace lbm {
namespace colk {

using namespace dev;

// Four values of one near-wall node, loaded or stored by one instruction
// (two for FP64).
template <class C>
struct alignas(4 * sizeof(C)) Quad {
  C x, y, z, w;
};

// Nontemporal store and load of a Quad (streamed data: written once, read
// once much later).
template <class C>
__device__ inline void store_quad_nt(Quad<C>* dst, const Quad<C>& v) {
  typedef C V4 __attribute__((ext_vector_type(4)));
  V4 x;
  x.x = v.x;
  x.y = v.y;
  x.z = v.z;
  x.w = v.w;
  __builtin_nontemporal_store(x, reinterpret_cast<V4*>(dst));
}
template <class C>
__device__ inline Quad<C> load_quad_nt(const Quad<C>* src) {
  typedef C V4 __attribute__((ext_vector_type(4)));
  const V4 x = __builtin_nontemporal_load(reinterpret_cast<const V4*>(src));
  return Quad<C>{x.x, x.y, x.z, x.w};
}

template <class L, class S>
struct ColKern {
  using P = Pop<S>;
  using C = typename P::C;
  static constexpr int Q = L::Q;
  static constexpr bool SHIFT = P::shifted;

  // Pull of a regular slot p in segment `seg`; bit q of `bm` marks a missing
  // link (half-way bounce-back: the node's own population opp(q)).
  struct Deltas {
    int4 lo, hi;
  };
  __device__ static Deltas load_deltas(const ColArgs& a, int32_t seg) {
    const int4* d4 = reinterpret_cast<const int4*>(a.seg_delta) + 2 * int64_t(seg);
    return Deltas{d4[0], d4[1]};
  }
  //
  // A32: the state buffer has fewer than 2^32 bytes, so every access is the
  // buffer's scalar base plus a 32-bit byte offset (q npad + slot) size:
  // the plane term q npad size is wave-uniform, the slot term one add per
  // lateral offset, and a bounce-back link swaps in the own-slot offset of
  // plane opp(q) with one bit-select. Otherwise 64-bit element indices.
  // Byte offsets of the regular pull of slot p (A32 layout, see below).
  __device__ static void regular_offsets(const ColArgs& a, int32_t p, const Deltas& dd, uint32_t bm, uint32_t* off) {
    constexpr uint32_t SZ = sizeof(S);
    const int32_t d[NCOL] = {dd.lo.x, dd.lo.y, dd.lo.z, dd.lo.w, dd.hi.x, dd.hi.y, dd.hi.z, dd.hi.w};
    const uint32_t plane = uint32_t(a.npad) * SZ;
    const uint32_t own = uint32_t(p) * SZ;
    uint32_t t[NCOL];
#pragma unroll
    for (int o = 0; o < NCOL; ++o) t[o] = own + uint32_t(d[o]) * SZ;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int o = col_of(L::cx[q], L::cy[q]);
      // (wrapping arithmetic: the sums are exact byte offsets < 2^32)
      const uint32_t reg = (o < 0 ? own : t[o]) + (uint32_t(q) * plane - uint32_t(L::cz[q] * int32_t(SZ)));
      const uint32_t bb = own + uint32_t(L::opp[q]) * plane;
      const uint32_t mask = uint32_t(__builtin_amdgcn_sbfe(int32_t(bm), q, 1));  // 0 or ~0
      off[q] = (bb & mask) | (reg & ~mask);
    }
  }

  // Byte offsets of the pull of explicit row `row` (-1 = bounce back).
  __device__ static void irregular_offsets(const ColArgs& a, int32_t p, int32_t row, uint32_t* off) {
    constexpr uint32_t SZ = sizeof(S);
    int32_t src[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) src[q] = a.irr_src[int64_t(q) * a.n_irr + row];
    const uint32_t plane = uint32_t(a.npad) * SZ;
    const uint32_t own = uint32_t(p) * SZ;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint32_t mask = uint32_t(src[q] >> 31);
      const uint32_t reg = uint32_t(src[q]) * SZ + uint32_t(q) * plane;
      const uint32_t bb = own + uint32_t(L::opp[q]) * plane;
      off[q] = (bb & mask) | (reg & ~mask);
    }
  }

  template <bool NTL = false, bool A32 = false>
  __device__ static void pull_regular(const ColArgs& a, int32_t p, const Deltas& dd, uint32_t bm, C* F) {
    const int32_t d[NCOL] = {dd.lo.x, dd.lo.y, dd.lo.z, dd.lo.w, dd.hi.x, dd.hi.y, dd.hi.z, dd.hi.w};
    if constexpr (A32) {
      uint32_t off[Q];
      regular_offsets(a, p, dd, bm, off);
#pragma unroll
      for (int q = 0; q < Q; ++q) F[q] = ld_off<S, NTL>(a.fs_in, off[q]);
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int o = col_of(L::cx[q], L::cy[q]);
        const int32_t s = (o < 0 ? p : p + d[o]) - L::cz[q];
        const bool bb = (bm >> q) & 1u;
        const int64_t idx = int64_t(bb ? L::opp[q] : q) * a.npad + (bb ? p : s);
        if constexpr (NTL) F[q] = P::ld_nt(a.fs_in, idx);
        else F[q] = P::ld(a.fs_in, idx);
      }
    }
  }

  // Pull of an explicit row.
  // (The same cache hint as the regular pull: the compiler merges the two
  // branches' loads and would otherwise drop the hint.)
  template <bool NTL = false, bool A32 = false>
  __device__ static void pull_irregular(const ColArgs& a, int32_t p, int32_t row, C* F) {
    if constexpr (A32) {
      uint32_t off[Q];
      irregular_offsets(a, p, row, off);
#pragma unroll
      for (int q = 0; q < Q; ++q) F[q] = ld_off<S, NTL>(a.fs_in, off[q]);
      return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int32_t s = a.irr_src[int64_t(q) * a.n_irr + row];
      const int32_t slot = s >= 0 ? s : p;
      const int qq = s >= 0 ? q : L::opp[q];
      if constexpr (NTL) F[q] = P::ld_nt(a.fs_in, int64_t(qq) * a.npad + slot);
      else F[q] = P::ld(a.fs_in, int64_t(qq) * a.npad + slot);
    }
  }

  // Pull of any slot, metadata read per lane.
  __device__ static void pull_any(const ColArgs& a, int32_t p, C* F) {
    const ChunkMeta m = a.meta[p / CHUNK];
    const int l = p & (CHUNK - 1);
    if ((m.irr_mask >> l) & 1u) {
      pull_irregular(a, p, m.irr_base + rank_below(m.irr_mask, l), F);
    } else {
      const uint32_t bm = ((m.bb_mask >> l) & 1u) ? a.bbm[m.bb_base + rank_below(m.bb_mask, l)] : 0u;
      pull_regular(a, p, load_deltas(a, segment_of(m, l)), bm, F);
    }
  }

  __device__ static void load_direct(const ColArgs& a, int32_t p, C* F) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double f = a.f_direct ? a.f_direct[int64_t(q) * a.npad + p] : L::w[q];
      F[q] = SHIFT ? C(f - L::w[q]) : C(f);
    }
  }

  __device__ static void write_direct(const ColArgs& a, int32_t p, const C* F) {
#pragma unroll
    for (int q = 0; q < Q; ++q) a.f_out[int64_t(q) * a.npad + p] = SHIFT ? double(F[q]) + L::w[q] : double(F[q]);
  }

  // Collide and store; near-wall node i (>= 0) also keeps the density
  // (or its deviation) and velocity of this collision for the next step's
  // wall shear stress.
  // KRES: planes [0, KRES) take plain accesses whatever the step's hint
  // (they stay cache-resident in both state buffers).
  template <bool FORCE, bool NT, bool A32 = false, int KRES = 0>
  __device__ static void collide_store(const ColArgs& a, int32_t p, C* F, int32_t i) {
    C drho, ux, uy, uz;
    const C frc[3] = {C(a.force[0]), C(a.force[1]), C(a.force[2])};
    const C alpha = (FORCE && a.porosity) ? C(a.porosity[p]) : C(0);
    collide_ext<L, C, SHIFT, FORCE>(F, C(a.omega), frc, alpha, false, drho, ux, uy, uz);
    if (i >= 0) static_cast<Quad<C>*>(a.nw_rho_u)[i] = Quad<C>{drho, ux, uy, uz};
    // NT: nontemporal stores (a compile-time choice: a runtime branch
    // between two otherwise identical stores is merged and loses the hint).
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if constexpr (A32) {
        // plane base (scalar) + own-slot offset (one vector register for all q)
        if (q < KRES)
          st_off<S, false>(static_cast<char*>(a.fs_out) + q * a.npad * int64_t(sizeof(S)),
                           uint32_t(p) * uint32_t(sizeof(S)), F[q]);
        else
          st_off<S, NT>(static_cast<char*>(a.fs_out) + q * a.npad * int64_t(sizeof(S)), uint32_t(p) * uint32_t(sizeof(S)),
                        F[q]);
      } else {
        if constexpr (NT) P::st_nt(a.fs_out, int64_t(q) * a.npad + p, F[q]);
        else P::st(a.fs_out, int64_t(q) * a.npad + p, F[q]);
      }
    }
  }

  // Wall shear stress of near-wall node i from its pulled,
  // boundary-treated populations F:
  //   tau_i = C sum_a (f_a - feq_a) (c_a.n) (c_ai - (c_a.n) n_i)
  // (Simulation.cpp:775-797 with its inner sums written out). feq is the
  // equilibrium of the collision that produced this step's input.
  // With Pi = sum_a (f_a - feq_a) c_a c_a^T this is tau = C (Pi n - (n.Pi n) n).
  // The fourth-order weight sums of D3Q19 and D3Q27 are isotropic, so the
  // second-order equilibrium has the exact second moment rho/3 I + rho u u^T
  // (shifted, w (drho + rho P): drho/3 I + rho u u^T), and Pi takes 42
  // additions over F instead of a rebuilt f_eq per direction.
  struct Eq {
    C s, rho, ux, uy, uz;
  };

  // Inputs of the stress of near-wall node i: the density (deviation) and
  // velocity that collide_store kept, and the wall normal, both in the
  // arithmetic precision (two vector loads).
  struct NwIn {
    C s, ux, uy, uz, nx, ny, nz;
  };
  __device__ static NwIn load_near_wall(const ColArgs& a, int32_t i) {
    const Quad<C> e = static_cast<const Quad<C>*>(a.nw_rho_u)[i];
    const Quad<C> n = static_cast<const Quad<C>*>(a.nw_normal_c)[i];
    return NwIn{e.x, e.y, e.z, e.w, n.x, n.y, n.z};
  }

  // The same density and velocity recovered from the node's own
  // post-collision populations, which BGK and the Guo source leave with the
  // same density and with momentum rho u (1 - alpha/2) + F/2 (used by the
  // adjoint, which differentiates through this map).
  template <bool FORCE>
  __device__ static Eq wss_equilibrium(const ColArgs& a, int32_t p) {
    C s = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const C v = P::ld(a.fs_in, int64_t(q) * a.npad + p);
      s += v;
      if (L::cx[q] > 0) jx += v;
      if (L::cx[q] < 0) jx -= v;
      if (L::cy[q] > 0) jy += v;
      if (L::cy[q] < 0) jy -= v;
      if (L::cz[q] > 0) jz += v;
      if (L::cz[q] < 0) jz -= v;
    }
    const C rho = SHIFT ? C(1) + s : s;
    C den = rho;
    if (FORCE) {
      const C alpha = a.porosity ? C(a.porosity[p]) : C(0);
      jx -= C(0.5) * C(a.force[0]);
      jy -= C(0.5) * C(a.force[1]);
      jz -= C(0.5) * C(a.force[2]);
      den = rho * (C(1) - C(0.5) * alpha);
    }
    return Eq{s, rho, jx / den, jy / den, jz / den};
  }

  // The stress of this step goes to ring slot a.ring_slot (4 values in the
  // arithmetic precision, plain stores, no dependent loads);
  // launch_wss_flush folds the ring into the FP64 accumulators.
  __device__ static void wss_node(const ColArgs& a, int32_t i, const NwIn& e, const C* F, double wt) {
#pragma clang fp contract(fast)
    Sum<C> Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      Pxx.acc(cx * cx, F[q]);
      Pyy.acc(cy * cy, F[q]);
      Pzz.acc(cz * cz, F[q]);
      Pxy.acc(cx * cy, F[q]);
      Pxz.acc(cx * cz, F[q]);
      Pyz.acc(cy * cz, F[q]);
    }
    C xx = Pxx.v, yy = Pyy.v, zz = Pzz.v, xy = Pxy.v, xz = Pxz.v, yz = Pyz.v;
    const C ux = e.ux, uy = e.uy, uz = e.uz, rho = SHIFT ? C(1) + e.s : e.s;
    const C iso = (SHIFT ? e.s : rho) * C(1.0 / 3.0);
    const C rx = rho * ux, ry = rho * uy;
    xx -= iso + rx * ux;
    yy -= iso + ry * uy;
    zz -= iso + rho * uz * uz;
    xy -= rx * uy;
    xz -= rx * uz;
    yz -= ry * uz;
    const int64_t m = a.n_nw;
    const C nx = e.nx, ny = e.ny, nz = e.nz;
    const C px = xx * nx + xy * ny + xz * nz;
    const C py = xy * nx + yy * ny + yz * nz;
    const C pz = xz * nx + yz * ny + zz * nz;
    const C npn = px * nx + py * ny + pz * nz;
    const C tx = px - npn * nx, ty = py - npn * ny, tz = pz - npn * nz;
    const C wc = C(a.wss_C);
    const C t0 = wc * tx, t1 = wc * ty, t2 = wc * tz;
    const C W = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    // Streamed past the caches: the ring is read back only by the flush,
    // and must not displace the state a cache-resident step relies on.
    store_quad_nt(static_cast<Quad<C>*>(a.nw_ring) + int64_t(a.ring_slot) * m + i, Quad<C>{t0, t1, t2, W});
    // Objective: weighted sum of |tau|^2 over the chosen wall segment (wt
    // was loaded with the pull; zero outside the segment or the adjoint).
    if (wt != 0.0) a.obj_acc[i] += wt * double(W) * double(W);
  }

  __device__ static bool is_boundary_kind(const ColArgs& a, uint32_t kind) {
    return (kind == 2 && a.inlet_mode != 2) || (kind == 3 && a.outlet_mode != 2);
  }

  // Slot p = lane l of a chunk whose metadata m is wave-uniform.
  //
  // Load order matters: the wait counter retires loads in issue order, and
  // after a branch that issued loads the compiler waits for everything
  // issued before the value it needs. So the stress inputs (raw loads, no
  // arithmetic on them until the stress) go first, then the pull-address
  // inputs (segment deltas, bounce-back mask; lanes without a mask read
  // entry 0, the array holds at least one): waiting for the latter costs
  // nothing more, and the pull is issued one memory latency after the
  // chunk metadata.
  template <int MODE, bool FORCE, bool NT, bool NTL, bool A32 = false, int KRES = 0>
  __device__ static void main_slot(const ColArgs& a, int32_t p, int l, const ChunkMeta& m) {
    C F[Q];
    const bool irr = (m.irr_mask >> l) & 1u;
    const int32_t row = irr ? m.irr_base + rank_below(m.irr_mask, l) : -1;
    if (MODE != KMODE_INIT && irr && is_boundary_kind(a, a.irr_kind[row] & ROW_KIND_MASK)) return;
    const int32_t i = ((m.nw_mask >> l) & 1u) ? m.nw_base + rank_below(m.nw_mask, l) : -1;
    const bool nw = MODE == KMODE_STEP && a.wss_on && i >= 0;
    NwIn e{};
    if (nw) e = load_near_wall(a, i);
    __builtin_amdgcn_sched_barrier(0);
    const Deltas dd = load_deltas(a, segment_of(m, l));
    const bool bbl = (m.bb_mask >> l) & 1u;
    const uint32_t bm_raw = a.bbm[bbl ? m.bb_base + rank_below(m.bb_mask, l) : 0];
    if (MODE == KMODE_INIT) {
      load_direct(a, p, F);
    } else if constexpr (A32) {
      // Offsets from either branch, then one set of scalar-base loads (loads
      // written in both branches would be merged with 64-bit addresses).
      uint32_t off[Q];
      if (irr) irregular_offsets(a, p, row, off);
      else regular_offsets(a, p, dd, bbl ? bm_raw : 0u, off);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q < KRES) F[q] = ld_off<S, false>(a.fs_in, off[q]);
        else F[q] = ld_off<S, NTL>(a.fs_in, off[q]);
      }
    } else if (irr) {
      pull_irregular<NTL>(a, p, row, F);
    } else {
      pull_regular<NTL>(a, p, dd, bbl ? bm_raw : 0u, F);
    }
    if (nw) wss_node(a, i, e, F, a.obj_acc ? a.obj_w[i] : 0.0);
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    if constexpr (MODE == KMODE_TAPE) store_tape(a, p, F);
    collide_store<FORCE, NT, A32, KRES>(a, p, F, i);
    if constexpr (MODE == KMODE_TAPE) store_sparse(a, p, F);
  }

  // KMODE_TAPE: the post-collision populations F of a flagged slot also go
  // to the sparse state copy (the few slots whose values the reverse step
  // still reads from the state: pull sources and own slots of objective
  // and inlet/outlet nodes).
  __device__ static void store_sparse(const ColArgs& a, int32_t p, const C* F) {
    if (!a.spec_out || !a.spec_flag[p]) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) P::st_nt(a.spec_out, int64_t(q) * a.npad + p, F[q]);
  }

  // Pull source slot of every direction of slot p (the node's own slot for a
  // missing link), decoded from the layout as the pull does.
  __device__ static void pull_slots(const ColArgs& a, int32_t p, int32_t* src) {
    const ChunkMeta m = a.meta[p / CHUNK];
    const int l = p & (CHUNK - 1);
    if ((m.irr_mask >> l) & 1u) {
      const int32_t row = m.irr_base + rank_below(m.irr_mask, l);
      for (int q = 0; q < Q; ++q) {
        const int32_t s = a.irr_src[int64_t(q) * a.n_irr + row];
        src[q] = s >= 0 ? s : p;
      }
      return;
    }
    const uint32_t bm = ((m.bb_mask >> l) & 1u) ? a.bbm[m.bb_base + rank_below(m.bb_mask, l)] : 0u;
    const Deltas dd = load_deltas(a, segment_of(m, l));
    const int32_t d[NCOL] = {dd.lo.x, dd.lo.y, dd.lo.z, dd.lo.w, dd.hi.x, dd.hi.y, dd.hi.z, dd.hi.w};
    for (int q = 0; q < Q; ++q) {
      const int o = col_of(L::cx[q], L::cy[q]);
      src[q] = ((bm >> q) & 1u) ? p : (o < 0 ? p : p + d[o]) - L::cz[q];
    }
  }

  // Moments of slot p's pulled populations for the adjoint's collision
  // transpose (read back once, much later: streamed past the caches).
  __device__ static void store_tape(const ColArgs& a, int32_t p, const C* F) {
    C s, jx, jy, jz;
    moments<L, C>(F, s, jx, jy, jz);
    C* t = static_cast<C*>(a.tape) + p;
    __builtin_nontemporal_store(s, t);
    __builtin_nontemporal_store(jx, t + a.npad);
    __builtin_nontemporal_store(jy, t + 2 * a.npad);
    __builtin_nontemporal_store(jz, t + 3 * a.npad);
  }

  // Unrolled to the fixed mode bound so the basis loads issue together.
  __device__ static C inlet_speed(const ColArgs& a, int32_t i) {
    double u = 0.0;
#pragma unroll
    for (int m = 0; m < MAX_MODES; ++m)
      if (m < a.n_modes) u += a.coef[m] * a.inlet_basis[int64_t(m) * a.n_inlet + i];
    return C(u);
  }

  // Non-equilibrium extrapolation in place on the interior neighbour's
  // post-stream populations G: G <- feq(rho_b, u_b) + (G - feq(rho_G, u_G)).
  // rho_b <= 0 takes rho_G; use_G_velocity takes u_b = u_G.
  __device__ static void nee_inplace(C* G, C rho_b, C bx, C by, C bz, bool use_G_velocity) {
#pragma clang fp contract(fast)
    C s, jx, jy, jz;
    moments<L, C>(G, s, jx, jy, jz);
    const C rg = SHIFT ? C(1) + s : s;
    const C gx = jx / rg, gy = jy / rg, gz = jz / rg;
    if (use_G_velocity) {
      bx = gx;
      by = gy;
      bz = gz;
    }
    if (rho_b <= C(0)) rho_b = rg;
    const C usq_b = bx * bx + by * by + bz * bz, usq_g = gx * gx + gy * gy + gz * gz;
    const C db = SHIFT ? rho_b - C(1) : C(0), dg = SHIFT ? s : C(0);
#pragma unroll
    for (int q = 0; q < Q; ++q)
      G[q] = geq_of<L, C, SHIFT>(q, db, rho_b, bx, by, bz, usq_b) +
             (G[q] - geq_of<L, C, SHIFT>(q, dg, rg, gx, gy, gz, usq_g));
  }

  // Inlet/outlet node p (always an explicit row).
  template <int MODE, bool FORCE, bool NT>
  __device__ static void boundary_slot(const ColArgs& a, int32_t p, int32_t bi) {
    const ChunkMeta m = a.meta[p / CHUNK];
    const int l = p & (CHUNK - 1);
    if (!((m.irr_mask >> l) & 1u)) return;
    const int32_t row = m.irr_base + rank_below(m.irr_mask, l);
    const uint32_t kind = a.irr_kind[row] & ROW_KIND_MASK;
    const int32_t u = a.irr_bcn[row];
    const int32_t i = ((m.nw_mask >> l) & 1u) ? m.nw_base + rank_below(m.nw_mask, l) : -1;
    const bool nw = MODE == KMODE_STEP && a.wss_on && i >= 0;
    NwIn e{};
    double wt = 0.0;
    if (nw) {
      e = load_near_wall(a, i);
      if (a.obj_acc) wt = a.obj_w[i];
    }
    C F[Q];
    if (kind == 2) {
      if (a.inlet_mode == INLET_NEE && u >= 0) {
        pull_any(a, u, F);
        const C U = inlet_speed(a, a.irr_inlet[row]);
        nee_inplace(F, C(-1), -C(a.inlet_normal[0]) * U, -C(a.inlet_normal[1]) * U, -C(a.inlet_normal[2]) * U,
                    false);
      } else {
        pull_irregular(a, p, row, F);
        if (a.inlet_mode == INLET_ZOUHE) {
          if constexpr (Q == 19) zouhe_ref<C, SHIFT>(F, inlet_speed(a, a.irr_inlet[row]), a.inlet_normal);
        }
      }
    } else if (u >= 0) {  // outlet with a node behind it
      pull_any(a, u, F);
      if (a.outlet_mode == OUTLET_REF) {
        // Simulation::ApplyOutletBc: every population of the node behind,
        // after that node's own inlet rule (Zou-He runs first).
        const ChunkMeta mu = a.meta[u / CHUNK];
        const int lu = u & (CHUNK - 1);
        if (((mu.irr_mask >> lu) & 1u) && a.inlet_mode == INLET_ZOUHE) {
          const int32_t urow = mu.irr_base + rank_below(mu.irr_mask, lu);
          if ((a.irr_kind[urow] & ROW_KIND_MASK) == 2) {
            if constexpr (Q == 19) zouhe_ref<C, SHIFT>(F, inlet_speed(a, a.irr_inlet[urow]), a.inlet_normal);
          }
        }
      } else {
        nee_inplace(F, C(a.outlet_rho), C(0), C(0), C(0), true);
      }
    } else {
      pull_irregular(a, p, row, F);
    }
    // The stress follows every boundary rule, as in the reference order.
    if (nw) wss_node(a, i, e, F, wt);
    if (MODE == KMODE_STEP && a.obj_rho) {  // objective: density sums for the pressure drop
      C s = C(0);
#pragma unroll
      for (int q = 0; q < Q; ++q) s += F[q];
      a.obj_rho[bi] += double(SHIFT ? C(1) + s : s);
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<FORCE, NT>(a, p, F, i);
    if constexpr (MODE == KMODE_TAPE) store_sparse(a, p, F);
  }
};

}  // namespace colk
}  // namespace lbm
