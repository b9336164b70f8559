// Column-offset kernel family (see columns.h for the layout).
//
// One launch per reference iteration:
//   leading blocks  inlet/outlet nodes: pull, inlet/outlet rule, collide,
//                   store (a short dependent-load chain; dispatched first so
//                   its latency hides under the rest of the launch);
//   other blocks    every other slot: pull (regular slots from 8 column bases
//                   + 16-bit offsets, irregular rows from their int32 sources
//                   with half-way bounce-back on missing links), collide,
//                   store; near-wall rows also accumulate the wall shear
//                   stress (Simulation::ComputeWallShearStress,
//                   Simulation.cpp:767-799) from the pulled populations and
//                   the density/velocity they stored at the previous collision.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../lattice.h"
#include "collision.h"
#include "columns.h"
#include "kernels.h"

namespace lbm {

namespace {

using namespace dev;

template <class L, class S>
struct ColKern {
  using P = Pop<S>;
  using C = typename P::C;
  static constexpr int Q = L::Q;
  static constexpr bool SHIFT = P::shifted;

  // Irregular row index of slot p, or -1.
  __device__ static int32_t irregular_row(const ColArgs& a, int32_t p) {
    if (a.idx16[p] != IRREGULAR) return -1;
    return int32_t(uint32_t(a.idx16[a.npad + p]) | (uint32_t(a.idx16[2 * a.npad + p]) << 16));
  }

  // Pull of a regular slot; `chunk` is wave-uniform in the main blocks
  // (scalar base loads).
  __device__ static void pull_regular(const ColArgs& a, int32_t p, int32_t chunk, C* F) {
    int32_t col[NCOL];
#pragma unroll
    for (int o = 0; o < NCOL; ++o) col[o] = a.cbase[o * a.nchunk + chunk] + int32_t(a.idx16[o * a.npad + p]);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int o = col_of(L::cx[q], L::cy[q]);
      const int32_t s = (o < 0 ? p : col[o]) - L::cz[q];
      F[q] = P::ld(a.fs_in, int64_t(q) * a.npad + s);
    }
  }

  // Pull of an irregular row: explicit sources, bounce back on missing links.
  __device__ static void pull_irregular(const ColArgs& a, int32_t p, int32_t row, C* F) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int32_t s = a.irr_src[int64_t(q) * a.n_irr + row];
      const int32_t slot = s >= 0 ? s : p;
      const int qq = s >= 0 ? q : L::opp[q];
      F[q] = P::ld(a.fs_in, int64_t(qq) * a.npad + slot);
    }
  }

  __device__ static void pull_any(const ColArgs& a, int32_t p, C* F) {
    const int32_t row = irregular_row(a, p);
    if (row >= 0) pull_irregular(a, p, row, F);
    else pull_regular(a, p, p / CHUNK, F);
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

  // Collide and store; a near-wall row keeps its density/velocity for the
  // next step's stress.
  template <bool FORCE>
  __device__ static void collide_store(const ColArgs& a, int32_t p, C* F, int32_t nw_row) {
    C drho, ux, uy, uz;
    const C frc[3] = {C(a.force[0]), C(a.force[1]), C(a.force[2])};
    const C alpha = (FORCE && a.porosity) ? C(a.porosity[p]) : C(0);
    collide_ext<L, C, SHIFT, FORCE>(F, C(a.omega), frc, alpha, false, drho, ux, uy, uz);
    if (nw_row >= 0) {
      const int64_t m = a.n_irr;
      a.irr_rho_u[nw_row] = double(drho);
      a.irr_rho_u[m + nw_row] = double(ux);
      a.irr_rho_u[2 * m + nw_row] = double(uy);
      a.irr_rho_u[3 * m + nw_row] = double(uz);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) P::st(a.fs_out, int64_t(q) * a.npad + p, F[q]);
  }

  // tau_i = C sum_a (f_a - feq_a) (c_a.n) (c_ai - (c_a.n) n_i), the reference
  // expression (Simulation.cpp:775-797) with its inner sums written out.
  __device__ static void wss_row(const ColArgs& a, int32_t row, const C* F) {
    const int64_t m = a.n_irr;
    const double n[3] = {a.irr_normal[row], a.irr_normal[m + row], a.irr_normal[2 * m + row]};
    const C drho = C(a.irr_rho_u[row]), ux = C(a.irr_rho_u[m + row]), uy = C(a.irr_rho_u[2 * m + row]),
            uz = C(a.irr_rho_u[3 * m + row]);
    const C rho = SHIFT ? C(1) + drho : drho;
    const C usq = ux * ux + uy * uy + uz * uz;
    double tau[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int al = 0; al < Q; ++al) {
      const double fne = double(F[al] - geq_of<L, C, SHIFT>(al, SHIFT ? drho : C(0), rho, ux, uy, uz, usq));
      const double cn = L::cx[al] * n[0] + L::cy[al] * n[1] + L::cz[al] * n[2];
      tau[0] += fne * cn * (L::cx[al] - cn * n[0]);
      tau[1] += fne * cn * (L::cy[al] - cn * n[1]);
      tau[2] += fne * cn * (L::cz[al] - cn * n[2]);
    }
    const double t0 = a.wss_C * tau[0], t1 = a.wss_C * tau[1], t2 = a.wss_C * tau[2];
    const double W = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    a.irr_wss[row] = W;
    a.irr_wss[m + row] += t0;
    a.irr_wss[2 * m + row] += t1;
    a.irr_wss[3 * m + row] += t2;
    a.irr_wss[4 * m + row] += W;
  }

  __device__ static bool is_boundary_kind(const ColArgs& a, uint32_t kind) {
    return (kind == 2 && a.inlet_mode != 2) || (kind == 3 && a.outlet_mode != 2);
  }

  template <int MODE, bool FORCE>
  __device__ static void main_slot(const ColArgs& a, int32_t p, int32_t chunk) {
    C F[Q];
    const int32_t row = irregular_row(a, p);
    uint32_t kf = 0;
    if (row >= 0) {
      kf = a.irr_kind[row];
      if (MODE != KMODE_INIT && is_boundary_kind(a, kf & ROW_KIND_MASK)) return;  // leading blocks own it
    }
    if (MODE == KMODE_INIT) {
      load_direct(a, p, F);
    } else if (row >= 0) {
      pull_irregular(a, p, row, F);
      if (MODE == KMODE_STEP && (kf & ROW_NEARWALL) && a.wss_on) wss_row(a, row, F);
    } else {
      pull_regular(a, p, chunk, F);
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<FORCE>(a, p, F, (row >= 0 && (kf & ROW_NEARWALL)) ? row : -1);
  }

  __device__ static C inlet_speed(const ColArgs& a, int32_t i) {
    double u = 0.0;
    for (int m = 0; m < a.n_modes; ++m) u += a.coef[m] * a.inlet_basis[int64_t(m) * a.n_inlet + i];
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

  // Inlet/outlet node p (always an irregular row).
  template <int MODE, bool FORCE>
  __device__ static void boundary_slot(const ColArgs& a, int32_t p) {
    const int32_t row = irregular_row(a, p);
    if (row < 0) return;
    const uint32_t kf = a.irr_kind[row];
    const uint32_t kind = kf & ROW_KIND_MASK;
    const int32_t u = a.irr_bcn[row];
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
        const int32_t urow = irregular_row(a, u);
        if (urow >= 0 && (a.irr_kind[urow] & ROW_KIND_MASK) == 2 && a.inlet_mode == INLET_ZOUHE) {
          if constexpr (Q == 19) zouhe_ref<C, SHIFT>(F, inlet_speed(a, a.irr_inlet[urow]), a.inlet_normal);
        }
      } else {
        nee_inplace(F, C(a.outlet_rho), C(0), C(0), C(0), true);
      }
    } else {
      pull_irregular(a, p, row, F);
    }
    // The stress follows every boundary rule, as in the reference order.
    if (MODE == KMODE_STEP && (kf & ROW_NEARWALL) && a.wss_on) wss_row(a, row, F);
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<FORCE>(a, p, F, (kf & ROW_NEARWALL) ? row : -1);
  }
};

template <class L, class S, int MODE, bool FORCE>
__global__ __launch_bounds__(BLOCK) void columns_kernel(ColArgs a) {
  const int32_t b = blockIdx.x;
  if (b < a.b_blocks) {
    const int32_t i = b * BLOCK + threadIdx.x;
    if (MODE != KMODE_INIT && i < a.n_bslots) ColKern<L, S>::template boundary_slot<MODE, FORCE>(a, a.bslots[i]);
    return;
  }
  int32_t lb = b - a.b_blocks;
  if (a.xcd_remap) {
    const int x = b & 7;
    lb = a.xcd_offset[x] + (b - a.xcd_first[x]) / 8;
  }
  const int32_t p = lb * BLOCK + threadIdx.x;
  // One wavefront covers one 64-slot chunk: its base loads are scalar.
  const int32_t chunk = __builtin_amdgcn_readfirstlane(p / CHUNK);
  if (p < a.n) ColKern<L, S>::template main_slot<MODE, FORCE>(a, p, chunk);
}

__global__ __launch_bounds__(BLOCK) void wss_stats_cols(ColArgs a, double* partials) {
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double s = 0.0, mx = 0.0;
  if (i < a.n_irr && (a.irr_kind[i] & ROW_NEARWALL)) {
    const double w = a.irr_wss[i];
    s = w * w;
    mx = w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    mx = fmax(mx, __shfl_xor(mx, off, 64));
  }
  __shared__ double red[BLOCK / 64][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv][0] = s;
    red[wv][1] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0, mm = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      ss += red[w][0];
      mm = fmax(mm, red[w][1]);
    }
    partials[2 * blockIdx.x] = ss;
    partials[2 * blockIdx.x + 1] = mm;
  }
}

template <class L, class S, int MODE, bool FORCE>
hipError_t launch_one(const ColArgs& a, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>(a.b_blocks + a.m_blocks);
  if (grid) hipLaunchKernelGGL((columns_kernel<L, S, MODE, FORCE>), dim3(grid), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

template <class L, class S>
hipError_t launch_mode(int kmode, bool force, const ColArgs& a, hipStream_t st) {
  switch (kmode) {
    case KMODE_STEP:
      return force ? launch_one<L, S, KMODE_STEP, true>(a, st) : launch_one<L, S, KMODE_STEP, false>(a, st);
    case KMODE_INIT:
      return force ? launch_one<L, S, KMODE_INIT, true>(a, st) : launch_one<L, S, KMODE_INIT, false>(a, st);
    case KMODE_MATERIALIZE:
      return launch_one<L, S, KMODE_MATERIALIZE, false>(a, st);
  }
  return hipErrorInvalidValue;
}

template <class L>
hipError_t launch_prec(int precision, int kmode, bool force, const ColArgs& a, hipStream_t st) {
  switch (precision) {
    case 0: return launch_mode<L, double>(kmode, force, a, st);
    case 1: return launch_mode<L, float>(kmode, force, a, st);
    case 2: return launch_mode<L, __half>(kmode, force, a, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_columns(int Q, int precision, int kmode, bool force, const ColArgs& a, hipStream_t st) {
  if (Q == 19) return launch_prec<D3Q19>(precision, kmode, force, a, st);
  if (Q == 27) return launch_prec<D3Q27>(precision, kmode, force, a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_columns_wss_stats(const ColArgs& a, double* partials, double* out2, hipStream_t st) {
  const int32_t nb = (a.n_irr + BLOCK - 1) / BLOCK;
  if (nb == 0) return hipMemsetAsync(out2, 0, 2 * sizeof(double), st);
  hipLaunchKernelGGL(wss_stats_cols, dim3(nb), dim3(BLOCK), 0, st, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_finish(partials, nb, 2, true, out2, st);
}

}  // namespace lbm
