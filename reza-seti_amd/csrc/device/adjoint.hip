// Discrete adjoint of the column-segment step (see adjoint.h for the
// recursion). Same grid shape as the forward launch: leading blocks take the
// inlet/outlet nodes, the rest every other slot, XCD-remapped.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "adjoint.h"
#include "colkern.h"

namespace lbm {

namespace {

using namespace dev;

// Moments of a per-direction vector v weighted by w_q:
//   A = sum w v,  B = sum w v c,  D = sum w v c c^T.
// For a collision with equilibrium feq_q = w_q rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2),
//   sum_q v_q feq_q = rho Psi(u),   Psi(u) = A + 3 B.u + 4.5 u.D.u - 1.5 A u.u,
// and the Guo source sums to (1 - omega/2) grad Psi(u) . F_tot.
template <class L, class C>
struct WMom {
  C A, Bx, By, Bz, Dxx, Dyy, Dzz, Dxy, Dxz, Dyz;

  __device__ void of(const C* v) {
#pragma clang fp contract(fast)
    A = Bx = By = Bz = Dxx = Dyy = Dzz = Dxy = Dxz = Dyz = C(0);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      const C t = C(L::w[q]) * v[q];
      const int x = L::cx[q], y = L::cy[q], z = L::cz[q];
      A += t;
      if (x) Bx += C(x) * t;
      if (y) By += C(y) * t;
      if (z) Bz += C(z) * t;
      if (x) Dxx += t;
      if (y) Dyy += t;
      if (z) Dzz += t;
      if (x && y) Dxy += C(x * y) * t;
      if (x && z) Dxz += C(x * z) * t;
      if (y && z) Dyz += C(y * z) * t;
    }
  }
  __device__ C psi(C ux, C uy, C uz) const {
#pragma clang fp contract(fast)
    const C uDu = Dxx * ux * ux + Dyy * uy * uy + Dzz * uz * uz +
                  C(2) * (Dxy * ux * uy + Dxz * ux * uz + Dyz * uy * uz);
    return A + C(3) * (Bx * ux + By * uy + Bz * uz) + C(4.5) * uDu - C(1.5) * A * (ux * ux + uy * uy + uz * uz);
  }
  // grad Psi(u) = 3 B + 9 D u - 3 A u
  __device__ void grad(C ux, C uy, C uz, C& gx, C& gy, C& gz) const {
#pragma clang fp contract(fast)
    gx = C(3) * Bx + C(9) * (Dxx * ux + Dxy * uy + Dxz * uz) - C(3) * A * ux;
    gy = C(3) * By + C(9) * (Dxy * ux + Dyy * uy + Dyz * uz) - C(3) * A * uy;
    gz = C(3) * Bz + C(9) * (Dxz * ux + Dyz * uy + Dzz * uz) - C(3) * A * uz;
  }
  // Hessian of Psi applied to v: (9 D - 3 A I) v
  __device__ void hess(C vx, C vy, C vz, C& hx, C& hy, C& hz) const {
#pragma clang fp contract(fast)
    hx = C(9) * (Dxx * vx + Dxy * vy + Dxz * vz) - C(3) * A * vx;
    hy = C(9) * (Dxy * vx + Dyy * vy + Dyz * vz) - C(3) * A * vy;
    hz = C(9) * (Dxz * vx + Dyz * vy + Dzz * vz) - C(3) * A * vz;
  }
};

template <class L, class S>
struct AdjKern {
  using K = colk::ColKern<L, S>;
  using C = typename K::C;
  static constexpr int Q = L::Q;
  static constexpr bool SHIFT = K::SHIFT;

  __device__ static C ald(const void* b, int64_t i) { return static_cast<const C*>(b)[i]; }
  __device__ static void ast(void* b, int64_t i, C v) { static_cast<C*>(b)[i] = v; }
  template <bool NT>
  __device__ static C ald_h(const void* b, int64_t i) {
    if constexpr (NT) return __builtin_nontemporal_load(static_cast<const C*>(b) + i);
    else return static_cast<const C*>(b)[i];
  }
  template <bool NT>
  __device__ static void ast_h(void* b, int64_t i, C v) {
    if constexpr (NT) __builtin_nontemporal_store(v, static_cast<C*>(b) + i);
    else static_cast<C*>(b)[i] = v;
  }

  // Transpose of the pull at slot p: population q of p went to the node
  // p + c_q, the pull source of p for direction opp(q); where that link is
  // missing it went to p itself as direction opp(q) (bounce-back).
  template <bool NTIN>
  __device__ static void rpull_regular(const AdjArgs& a, int32_t p, int32_t seg, uint32_t bm, C* lam) {
    const ColArgs& c = a.c;
    const int4* d4 = reinterpret_cast<const int4*>(c.seg_delta) + 2 * int64_t(seg);
    const int4 d0 = d4[0], d1 = d4[1];
    const int32_t d[NCOL] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int qo = L::opp[q];
      const int o = col_of(L::cx[qo], L::cy[qo]);
      const int32_t s = (o < 0 ? p : p + d[o]) - L::cz[qo];
      const bool bb = (bm >> qo) & 1u;
      lam[q] = ald_h<NTIN>(a.rho_in, int64_t(bb ? qo : q) * c.npad + (bb ? p : s));
    }
  }

  template <bool NTIN>
  __device__ static void rpull_irregular(const AdjArgs& a, int32_t p, int32_t row, C* lam) {
    const ColArgs& c = a.c;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int qo = L::opp[q];
      const int32_t s = c.irr_src[int64_t(qo) * c.n_irr + row];
      lam[q] = ald_h<NTIN>(a.rho_in, s >= 0 ? int64_t(q) * c.npad + s : int64_t(qo) * c.npad + p);
    }
  }

  // mu = C'(G)^T lam for f* = G + omega (feq - G) + (1 - omega/2) S(G),
  //   u = (j + F/2) / (rho k), k = 1 + alpha/2, F_tot = F - alpha rho u:
  //   mu_p = (1 - omega) lam_p + Phi_rho + Phi_u . (c_p / (rho k) - u / rho),
  //   Phi(rho, u) = omega rho Psi(u) + (1 - omega/2) grad Psi(u) . F_tot.
  // G enters through its moments only (s = sum G, j = sum c G); lam is
  // replaced by mu in place. Returns d(lam . f*)/d alpha at fixed G.
  __device__ static double collide_adjoint(const AdjArgs& a, int32_t p, C s, C jx, C jy, C jz, C* lam) {
#pragma clang fp contract(fast)
    const ColArgs& c = a.c;
    const C rho = SHIFT ? C(1) + s : s;
    const C alpha = c.porosity ? C(c.porosity[p]) : C(0);
    const C Fx = C(c.force[0]), Fy = C(c.force[1]), Fz = C(c.force[2]);
    const C k = C(1) + C(0.5) * alpha;
    const C irk = C(1) / (rho * k), ir = C(1) / rho;
    const C ux = (jx + C(0.5) * Fx) * irk, uy = (jy + C(0.5) * Fy) * irk, uz = (jz + C(0.5) * Fz) * irk;
    const C Ftx = Fx - alpha * rho * ux, Fty = Fy - alpha * rho * uy, Ftz = Fz - alpha * rho * uz;
    const C om = C(c.omega), kap = C(1) - C(0.5) * om;
    WMom<L, C> M;
    M.of(lam);
    const C psi = M.psi(ux, uy, uz);
    C gx, gy, gz, hx, hy, hz;
    M.grad(ux, uy, uz, gx, gy, gz);
    M.hess(Ftx, Fty, Ftz, hx, hy, hz);
    const C gu = gx * ux + gy * uy + gz * uz;
    const C phr = om * psi - kap * alpha * gu;
    const C f1 = (om - kap * alpha) * rho;
    const C pux = f1 * gx + kap * hx, puy = f1 * gy + kap * hy, puz = f1 * gz + kap * hz;
    const C pu_u = pux * ux + puy * uy + puz * uz;
    const C base = phr - pu_u * ir;
#pragma unroll
    for (int q = 0; q < Q; ++q)
      lam[q] = (C(1) - om) * lam[q] + base + cdot<L, C>(q, pux, puy, puz) * irk;
    return double(-pu_u / (C(2) * k) - kap * rho * gu);
  }

  // Raw stress sum_a (G_a - feq_a) K_a (K_a = (c_a.n)(c_a - (c_a.n) n)) of
  // near-wall node i from its recomputed populations G and the equilibrium e.
  __device__ static void tau_raw(const ColArgs& c, int32_t i, const typename K::Eq& e, const C* G, C& tx, C& ty,
                                 C& tz) {
#pragma clang fp contract(fast)
    const C usq = e.ux * e.ux + e.uy * e.uy + e.uz * e.uz;
    const int64_t m = c.n_nw;
    const C nx = C(c.nw_normal[i]), ny = C(c.nw_normal[m + i]), nz = C(c.nw_normal[2 * m + i]);
    tx = ty = tz = C(0);
#pragma unroll
    for (int al = 0; al < Q; ++al) {
      const C fne = G[al] - geq_of<L, C, SHIFT>(al, SHIFT ? e.s : C(0), e.rho, e.ux, e.uy, e.uz, usq);
      const C cn = C(L::cx[al]) * nx + C(L::cy[al]) * ny + C(L::cz[al]) * nz;
      tx += fne * cn * (C(L::cx[al]) - cn * nx);
      ty += fne * cn * (C(L::cy[al]) - cn * ny);
      tz += fne * cn * (C(L::cz[al]) - cn * nz);
    }
  }

  // Stress term of the objective at near-wall node i (slot p), given its raw
  // stress (tx, ty, tz) and the equilibrium e recovered from the node's own
  // input populations: adds dJ_t/dG_t to mu, stores dJ_t/dh(p) (through that
  // recovery) for the next reverse step, and returns dJ_t/dalpha(p) of the
  // recovery. The moments of zeta = -dJ_t/dG are accumulated on the fly.
  __device__ static double wss_adjoint(const AdjArgs& a, int32_t p, int32_t i, const typename K::Eq& e, C tx, C ty,
                                       C tz, C* mu) {
#pragma clang fp contract(fast)
    const ColArgs& c = a.c;
    const int64_t m = c.n_nw;
    const C nx = C(c.nw_normal[i]), ny = C(c.nw_normal[m + i]), nz = C(c.nw_normal[2 * m + i]);
    // J_t = w C^2 |tau_raw|^2 / T  ->  dJ_t/dG_a = 2 w C^2 / T (tau_raw . K_a)
    const C coef = C(2.0 * a.obj_w[i] * a.inv_T * c.wss_C * c.wss_C);
    WMom<L, C> M;
    M.A = M.Bx = M.By = M.Bz = M.Dxx = M.Dyy = M.Dzz = M.Dxy = M.Dxz = M.Dyz = C(0);
#pragma unroll
    for (int al = 0; al < Q; ++al) {
      const int x = L::cx[al], y = L::cy[al], z = L::cz[al];
      const C cn = C(x) * nx + C(y) * ny + C(z) * nz;
      const C eta = coef * cn * ((C(x) - cn * nx) * tx + (C(y) - cn * ny) * ty + (C(z) - cn * nz) * tz);
      mu[al] += eta;
      const C t = -C(L::w[al]) * eta;  // w * zeta
      M.A += t;
      if (x) M.Bx += C(x) * t;
      if (y) M.By += C(y) * t;
      if (z) M.Bz += C(z) * t;
      if (x) M.Dxx += t;
      if (y) M.Dyy += t;
      if (z) M.Dzz += t;
      if (x && y) M.Dxy += C(x * y) * t;
      if (x && z) M.Dxz += C(x * z) * t;
      if (y && z) M.Dyz += C(y * z) * t;
    }
    // d/dh_q of rho_h Psi_zeta(u_h), u_h = (j_h - F/2) / (rho_h k'), k' = 1 - alpha/2.
    const C alpha = c.porosity ? C(c.porosity[p]) : C(0);
    const C kp = C(1) - C(0.5) * alpha;
    const C psi = M.psi(e.ux, e.uy, e.uz);
    C gx, gy, gz;
    M.grad(e.ux, e.uy, e.uz, gx, gy, gz);
    const C gu = gx * e.ux + gy * e.uy + gz * e.uz;
#pragma unroll
    for (int q = 0; q < Q; ++q)
      a.own[int64_t(i) * Q + q] =
          double(psi + cdot<L, C>(q, gx, gy, gz) / kp - gu);
    return double(e.rho * gu / (C(2) * kp));
  }

  __device__ static void add_own(const AdjArgs& a, int32_t i, C* lam) {
#pragma unroll
    for (int q = 0; q < Q; ++q) lam[q] += C(a.own[int64_t(i) * Q + q]);
  }

  // The recomputed state is read once per reverse step: stream it past the
  // caches, which keep the adjoint field written by the previous launch.
  __device__ static void pull_main(const AdjArgs& a, int32_t p, bool irr, int32_t row, int32_t seg, uint32_t bm,
                                   C* G) {
    if (irr) K::template pull_irregular<true>(a.c, p, row, G);
    else K::template pull_regular<true>(a.c, p, K::load_deltas(a.c, seg), bm, G);
  }

  // om: the chunk's objective mask (wave-uniform).
  template <bool NTIN, bool NTOUT>
  __device__ static void main_slot(const AdjArgs& a, int32_t p, int l, const ChunkMeta& m, uint64_t om) {
    const ColArgs& c = a.c;
    const bool irr = (m.irr_mask >> l) & 1u;
    const int32_t row = irr ? m.irr_base + rank_below(m.irr_mask, l) : -1;
    if (a.mode == ADJ_STEP && irr && K::is_boundary_kind(c, c.irr_kind[row] & ROW_KIND_MASK)) return;
    const uint32_t bm = (!irr && ((m.bb_mask >> l) & 1u)) ? c.bbm[m.bb_base + rank_below(m.bb_mask, l)] : 0u;
    const int32_t seg = segment_of(m, l);
    const bool nw = (m.nw_mask >> l) & 1u;
    const int32_t i = nw ? m.nw_base + rank_below(m.nw_mask, l) : -1;
    const bool in_obj = (om >> l) & 1u;  // (objective nodes are near-wall nodes)
    // dJ/dalpha is read and written once per reverse step: streamed, so that
    // it does not take Infinity Cache room from the adjoint field.
    const double g_old = a.g_alpha ? __builtin_nontemporal_load(a.g_alpha + p) : 0.0;
    // Moments of the step's pre-collision populations G_t: from the tape of
    // the recomputation, or recomputed here.
    C s, jx, jy, jz;
    if (a.mode == ADJ_STEP && a.tape_in) {
      const C* t = static_cast<const C*>(a.tape_in) + p;
      s = __builtin_nontemporal_load(t);
      jx = __builtin_nontemporal_load(t + c.npad);
      jy = __builtin_nontemporal_load(t + 2 * c.npad);
      jz = __builtin_nontemporal_load(t + 3 * c.npad);
    } else {
      C G[Q];
      if (a.mode == ADJ_FINAL) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const double f = a.g0 ? a.g0[int64_t(q) * c.npad + p] : L::w[q];
          G[q] = SHIFT ? C(f - L::w[q]) : C(f);
        }
      } else {
        pull_main(a, p, irr, row, seg, bm, G);
      }
      moments<L, C>(G, s, jx, jy, jz);
    }
    // The objective's stress at this node, from the recomputed populations,
    // before the adjoint field is loaded: the two 19-value arrays are never
    // live together, which keeps this rare path within the register count
    // of the common one (and so the whole launch at more waves per SIMD).
    typename K::Eq e{};
    C tx = C(0), ty = C(0), tz = C(0);
    if (a.mode == ADJ_STEP && in_obj) {
      e = K::template wss_equilibrium<true>(c, p);
      C G[Q];
      pull_main(a, p, irr, row, seg, bm, G);
      tau_raw(c, i, e, G, tx, ty, tz);
    }
    C lam[Q];
    if (a.rho_in) {
      if (irr) rpull_irregular<NTIN>(a, p, row, lam);
      else rpull_regular<NTIN>(a, p, seg, bm, lam);
      if (in_obj) add_own(a, i, lam);
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) lam[q] = C(0);
    }
    double dal = collide_adjoint(a, p, s, jx, jy, jz, lam);  // lam -> mu
    if (a.mode == ADJ_STEP && in_obj) dal += wss_adjoint(a, p, i, e, tx, ty, tz, lam);
    if (a.g_alpha) __builtin_nontemporal_store(g_old + dal, a.g_alpha + p);
    if (a.mode == ADJ_FINAL) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) ast_h<NTOUT>(a.rho_out, int64_t(q) * c.npad + p, lam[q]);
  }

  // Post-rule populations of inlet/outlet node p (as the forward step).
  __device__ static void boundary_G(const ColArgs& c, int32_t p, int32_t row, uint32_t kind, int32_t u, bool nee,
                                    C* G) {
    if (nee) {
      K::pull_any(c, u, G);
      if (kind == 2) {
        const C U = K::inlet_speed(c, c.irr_inlet[row]);
        K::nee_inplace(G, C(-1), -C(c.inlet_normal[0]) * U, -C(c.inlet_normal[1]) * U, -C(c.inlet_normal[2]) * U,
                       false);
      } else {
        K::nee_inplace(G, C(c.outlet_rho), C(0), C(0), C(0), true);
      }
    } else {
      K::pull_irregular(c, p, row, G);
    }
  }

  // Inlet/outlet node p (an explicit row), index bi in the boundary list.
  __device__ static void boundary_slot(const AdjArgs& a, int32_t p, int32_t bi) {
    const ColArgs& c = a.c;
    const ChunkMeta m = c.meta[p / CHUNK];
    const int l = p & (CHUNK - 1);
    if (!((m.irr_mask >> l) & 1u)) return;
    const int32_t row = m.irr_base + rank_below(m.irr_mask, l);
    const uint32_t kind = c.irr_kind[row] & ROW_KIND_MASK;
    const int32_t u = c.irr_bcn[row];
    const bool nw = (m.nw_mask >> l) & 1u;
    const int32_t i = nw ? m.nw_base + rank_below(m.nw_mask, l) : -1;
    const bool in_obj = nw && a.obj_w && a.obj_w[i] != 0.0;
    const bool nee = u >= 0 && ((kind == 2 && c.inlet_mode == INLET_NEE) || (kind == 3 && c.outlet_mode == OUTLET_NEE));
    // Moments and (at an objective node) the stress of the post-rule
    // populations first, then the adjoint field (as in main_slot).
    C s, jx, jy, jz;
    typename K::Eq e{};
    C tx = C(0), ty = C(0), tz = C(0);
    {
      C G[Q];
      boundary_G(c, p, row, kind, u, nee, G);
      moments<L, C>(G, s, jx, jy, jz);
      if (in_obj) {
        e = K::template wss_equilibrium<true>(c, p);
        tau_raw(c, i, e, G, tx, ty, tz);
      }
    }
    C lam[Q];
    if (a.rho_in) {
      rpull_irregular<false>(a, p, row, lam);
      if (in_obj) add_own(a, i, lam);
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) lam[q] = C(0);
    }
    double dal = collide_adjoint(a, p, s, jx, jy, jz, lam);  // lam -> mu
    if (in_obj) dal += wss_adjoint(a, p, i, e, tx, ty, tz, lam);
    const C dp = kind == 2 ? C(a.dp_in) : C(-a.dp_out);
#pragma unroll
    for (int q = 0; q < Q; ++q) lam[q] += dp;
    if (a.g_alpha) a.g_alpha[p] += dal;
    if (nee) {
      const int64_t j = a.member_of[bi];
#pragma unroll
      for (int q = 0; q < Q; ++q) ast(a.mb, j * (Q + 1) + q, lam[q]);
      ast(a.mb, j * (Q + 1) + Q, kind == 2 ? K::inlet_speed(c, c.irr_inlet[row]) : C(0));
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) ast(a.rho_out, int64_t(q) * c.npad + p, nee ? C(0) : lam[q]);
  }

  // Upstream node y of extrapolated boundary nodes b: adds NEE_b'^T mu(b) to
  // rho(y), and dJ/dU_b for inlet nodes.
  //   inlet  G = feq(rho_R, U_b) + R - feq(rho_R, u_R):
  //          d/dR_p = mu_p + Psi(U_b) - Psi(u_R) - grad Psi(u_R).(c_p - u_R)
  //          d/dU_b = rho_R grad Psi(U_b) . (-n)        (U_b = -n U)
  //   outlet G = feq(rho_out, u_R) + R - feq(rho_R, u_R):
  //          d/dR_p = mu_p - Psi(u_R) + (rho_out / rho_R - 1) grad Psi(u_R).(c_p - u_R)
  // with Psi the w-weighted moments of mu(b).
  __device__ static void upstream(const AdjArgs& a, int32_t k) {
#pragma clang fp contract(fast)
    const ColArgs& c = a.c;
    const int32_t y = a.tgt[k];
    // rho(y) as kernel A wrote it, loaded first: its round trip overlaps the
    // member loads below instead of following them.
    C old[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) old[q] = ald(a.rho_out, int64_t(q) * c.npad + y);
    C s, jx, jy, jz;
    if (a.tape_in) {  // y is a main slot: its pulled moments are on the tape
      const C* t = static_cast<const C*>(a.tape_in) + y;
      s = t[0];
      jx = t[c.npad];
      jy = t[2 * c.npad];
      jz = t[3 * c.npad];
    } else {
      C R[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int32_t src = a.tgt_src[int64_t(k) * Q + q];
        R[q] = K::P::ld(c.fs_in, int64_t(src >= 0 ? q : L::opp[q]) * c.npad + (src >= 0 ? src : y));
      }
      moments<L, C>(R, s, jx, jy, jz);
    }
    const C rR = SHIFT ? C(1) + s : s;
    const C ux = jx / rR, uy = jy / rR, uz = jz / rR;
    C acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = C(0);
    for (int32_t j = a.tgt_off[k]; j < a.tgt_off[k + 1]; ++j) {
      const int32_t inl = a.tgt_inl[j];
      C mu[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) mu[q] = ald(a.mb, int64_t(j) * (Q + 1) + q);
      WMom<L, C> M;
      M.of(mu);
      const C psiR = M.psi(ux, uy, uz);
      C gx, gy, gz;
      M.grad(ux, uy, uz, gx, gy, gz);
      C cst, fac;
      if (inl >= 0) {
        const C U = ald(a.mb, int64_t(j) * (Q + 1) + Q);
        const C nx = C(c.inlet_normal[0]), ny = C(c.inlet_normal[1]), nz = C(c.inlet_normal[2]);
        const C bx = -nx * U, by = -ny * U, bz = -nz * U;
        cst = M.psi(bx, by, bz) - psiR;
        fac = C(-1);
        C hx, hy, hz;
        M.grad(bx, by, bz, hx, hy, hz);
        a.gamma[inl] = -double(rR) * double(hx * nx + hy * ny + hz * nz);
      } else {
        cst = -psiR;
        fac = C(c.outlet_rho) / rR - C(1);
      }
      const C gu = gx * ux + gy * uy + gz * uz;
#pragma unroll
      for (int q = 0; q < Q; ++q)
        acc[q] += mu[q] + cst + fac * (cdot<L, C>(q, gx, gy, gz) - gu);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) ast(a.rho_out, int64_t(q) * c.npad + y, old[q] + acc[q]);
  }
};

template <class L, class S>
__device__ inline void adjoint_body(const AdjArgs& a) {
  const ColArgs& c = a.c;
  const int32_t b = blockIdx.x;
  if (b < c.b_blocks) {
    const int32_t i = b * BLOCK + threadIdx.x;
    if (a.mode == ADJ_STEP && i < c.n_bslots) AdjKern<L, S>::boundary_slot(a, c.bslots[i], i);
    return;
  }
  const int32_t lb = main_block_of(c, b);
  if (lb < 0) return;
  const int32_t p = lb * BLOCK + threadIdx.x;
  const int32_t chunk = __builtin_amdgcn_readfirstlane(p / CHUNK);
  const ChunkMeta m = c.meta[chunk];
  const uint64_t om = a.obj_mask ? a.obj_mask[chunk] : 0;
  if (p >= c.n) return;
  if (a.rho_cached) AdjKern<L, S>::template main_slot<false, true>(a, p, p & (CHUNK - 1), m, om);
  else AdjKern<L, S>::template main_slot<true, false>(a, p, p & (CHUNK - 1), m, om);
}

template <class L, class S>
__global__ __launch_bounds__(BLOCK) void adjoint_kernel(AdjArgs a) {
  adjoint_body<L, S>(a);
}

#if LBM_ADJOINT_WAVES_PER_EU > 0
template <class L, class S>
__global__ __attribute__((amdgpu_waves_per_eu(LBM_ADJOINT_WAVES_PER_EU))) __launch_bounds__(BLOCK) void
adjoint_kernel_occ(AdjArgs a) {
  adjoint_body<L, S>(a);
}
#endif

// One wavefront per block: the few thousand upstream nodes spread over as
// many CUs as possible (the kernel is latency-bound).
constexpr int BBLOCK = 64;

template <class L, class S>
__global__ __launch_bounds__(BBLOCK) void adjoint_boundary_kernel(AdjArgs a) {
  const int32_t k = blockIdx.x * BBLOCK + threadIdx.x;
  if (k < a.n_tgt) AdjKern<L, S>::upstream(a, k);
}

__global__ __launch_bounds__(BLOCK) void coef_gradient_kernel(const double* gamma, const double* basis,
                                                              int32_t n_inlet, int32_t n_modes, double* out) {
  __shared__ double red[BLOCK / 64];
  const int32_t t = blockIdx.x;
  for (int m = 0; m < n_modes; ++m) {
    double s = 0.0;
    for (int32_t b = threadIdx.x; b < n_inlet; b += BLOCK)
      s += gamma[int64_t(t) * n_inlet + b] * basis[int64_t(m) * n_inlet + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int w = 0; w < BLOCK / 64; ++w) tot += red[w];
      out[int64_t(t) * n_modes + m] = tot;
    }
    __syncthreads();
  }
}

template <class L, class S>
__global__ __launch_bounds__(BLOCK) void spec_flag_kernel(ColArgs c, const int32_t* entries, int32_t n,
                                                          uint8_t* flag) {
  const int32_t k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= n) return;
  const int32_t p = entries[k];
  int32_t src[L::Q];
  colk::ColKern<L, S>::pull_slots(c, p, src);
  flag[p] = 1;
  for (int q = 0; q < L::Q; ++q) flag[src[q]] = 1;
}

template <class L, class S>
hipError_t launch_a(const AdjArgs& a0, hipStream_t st) {
  // The reverse step keeps one contiguous eighth of the slots per XCD. The
  // forward step's runs of 64 blocks map blocks to slots one to one as well
  // (the reverse step built with them passes every adjoint test and gives
  // bit-identical gradients on the 2.85M-node bifurcation), but they do not
  // shorten it there (reverse pass 0.919 s against 0.922 s).
  AdjArgs a = a0;
  a.c.xcd_run = 0;
  const unsigned grid = static_cast<unsigned>(a.c.b_blocks + main_grid_blocks(a.c));
  if (!grid) return hipSuccess;
#if LBM_ADJOINT_WAVES_PER_EU > 0
  if constexpr (!std::is_same<S, double>::value) {
    hipLaunchKernelGGL((adjoint_kernel_occ<L, S>), dim3(grid), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL((adjoint_kernel<L, S>), dim3(grid), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

template <class L, class S>
hipError_t launch_b(const AdjArgs& a, hipStream_t st) {
  if (a.n_tgt > 0)
    hipLaunchKernelGGL((adjoint_boundary_kernel<L, S>), dim3((a.n_tgt + BBLOCK - 1) / BBLOCK), dim3(BBLOCK), 0,
                       st, a);
  return hipGetLastError();
}

template <class L>
hipError_t dispatch(int precision, bool boundary, const AdjArgs& a, hipStream_t st) {
  switch (precision) {
    case 0: return boundary ? launch_b<L, double>(a, st) : launch_a<L, double>(a, st);
    case 1: return boundary ? launch_b<L, float>(a, st) : launch_a<L, float>(a, st);
    case 2: return boundary ? launch_b<L, __half>(a, st) : launch_a<L, __half>(a, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_spec_flags(int Q, int precision, const ColArgs& c, const int32_t* entries, int32_t n,
                             uint8_t* flag, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((n + BLOCK - 1) / BLOCK), blk(BLOCK);
  // The decoding reads only the layout, which is the same for every precision.
  (void)precision;
  if (Q == 19) hipLaunchKernelGGL((spec_flag_kernel<D3Q19, float>), grid, blk, 0, st, c, entries, n, flag);
  else if (Q == 27) hipLaunchKernelGGL((spec_flag_kernel<D3Q27, float>), grid, blk, 0, st, c, entries, n, flag);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_adjoint(int Q, int precision, const AdjArgs& a, hipStream_t st) {
  if (Q == 19) return dispatch<D3Q19>(precision, false, a, st);
  if (Q == 27) return dispatch<D3Q27>(precision, false, a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_adjoint_boundary(int Q, int precision, const AdjArgs& a, hipStream_t st) {
  if (Q == 19) return dispatch<D3Q19>(precision, true, a, st);
  if (Q == 27) return dispatch<D3Q27>(precision, true, a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_coef_gradient(const double* gamma, const double* basis, int32_t T, int32_t n_inlet,
                                int32_t n_modes, double* out, hipStream_t st) {
  if (T > 0 && n_modes > 0)
    hipLaunchKernelGGL(coef_gradient_kernel, dim3(T), dim3(BLOCK), 0, st, gamma, basis, n_inlet, n_modes, out);
  return hipGetLastError();
}

}  // namespace lbm
