// Column-offset kernel family (see columns.h for the layout).
//
// Per step two launches:
//   columns_kernel    every slot except inlet/outlet nodes: pull (regular
//                     slots from 8 column bases + 16-bit offsets, irregular
//                     slots from their int32 row with half-way bounce-back on
//                     missing links), collide, store. Lean on purpose: no
//                     boundary or WSS state, so its register budget allows
//                     full occupancy.
//   boundary_kernel   inlet/outlet nodes (pull, inlet/outlet rule, collide,
//                     store) and, on WSS steps, the near-wall stress
//                     (Simulation::ComputeWallShearStress, Simulation.cpp:767-799)
//                     from the same step's input populations.
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

  // Pull of a regular slot; `chunk` may be wave-uniform (scalar loads).
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

  // Pull of an irregular slot: explicit sources, bounce back on missing links.
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

  template <bool FORCE>
  __device__ static void collide_store(const ColArgs& a, int32_t p, C* F) {
    C drho, ux, uy, uz;
    const C frc[3] = {C(a.force[0]), C(a.force[1]), C(a.force[2])};
    const C alpha = (FORCE && a.porosity) ? C(a.porosity[p]) : C(0);
    collide_ext<L, C, SHIFT, FORCE>(F, C(a.omega), frc, alpha, false, drho, ux, uy, uz);
#pragma unroll
    for (int q = 0; q < Q; ++q) P::st(a.fs_out, int64_t(q) * a.npad + p, F[q]);
  }

  __device__ static bool is_boundary_kind(const ColArgs& a, uint32_t kind) {
    return (kind == 2 && a.inlet_mode != 2) || (kind == 3 && a.outlet_mode != 2);
  }

  template <int MODE, bool FORCE>
  __device__ static void main_slot(const ColArgs& a, int32_t p, int32_t chunk) {
    C F[Q];
    if (MODE == KMODE_INIT) {
      load_direct(a, p, F);
    } else {
      const int32_t row = irregular_row(a, p);
      if (row >= 0) {
        if (is_boundary_kind(a, a.irr_kind[row])) return;  // boundary_kernel owns it
        pull_irregular(a, p, row, F);
      } else {
        pull_regular(a, p, chunk, F);
      }
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<FORCE>(a, p, F);
  }

  __device__ static C inlet_speed(const ColArgs& a, int32_t i) {
    double u = 0.0;
    for (int m = 0; m < a.n_modes; ++m) u += a.coef[m] * a.inlet_basis[int64_t(m) * a.n_inlet + i];
    return C(u);
  }

  // Inlet/outlet node p (always irregular): pull, rule, collide, store.
  template <int MODE, bool FORCE>
  __device__ static void boundary_slot(const ColArgs& a, int32_t p) {
    const int32_t row = irregular_row(a, p);
    if (row < 0) return;
    const uint32_t kind = a.irr_kind[row];
    C F[Q];
    pull_irregular(a, p, row, F);
    const int32_t u = a.irr_bcn[row];
    if (kind == 2) {
      if (a.inlet_mode == INLET_ZOUHE) {
        if constexpr (Q == 19) zouhe_ref<C, SHIFT>(F, inlet_speed(a, a.irr_inlet[row]), a.inlet_normal);
      } else if (a.inlet_mode == INLET_NEE && u >= 0) {
        C G[Q];
        pull_any(a, u, G);
        const C U = inlet_speed(a, a.irr_inlet[row]);
        nee<L, C, SHIFT>(F, G, C(-1), -C(a.inlet_normal[0]) * U, -C(a.inlet_normal[1]) * U,
                         -C(a.inlet_normal[2]) * U, false);
      }
    } else if (kind == 3 && u >= 0) {
      C G[Q];
      pull_any(a, u, G);
      if (a.outlet_mode == OUTLET_REF) {
        const int32_t urow = irregular_row(a, u);
        if (urow >= 0 && a.irr_kind[urow] == 2 && a.inlet_mode == INLET_ZOUHE) {
          if constexpr (Q == 19) zouhe_ref<C, SHIFT>(G, inlet_speed(a, a.irr_inlet[urow]), a.inlet_normal);
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) F[q] = G[q];
      } else if (a.outlet_mode == OUTLET_NEE) {
        nee<L, C, SHIFT>(F, G, C(a.outlet_rho), C(0), C(0), C(0), true);
      }
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<FORCE>(a, p, F);
  }
};

__device__ inline int32_t logical_block(int32_t b, const int32_t* first, const int32_t* offset, int remap) {
  if (!remap) return b;
  const int x = b & 7;
  return offset[x] + (b - first[x]) / 8;
}

template <class L, class S, int MODE, bool FORCE>
__global__ __launch_bounds__(BLOCK) void columns_kernel(ColArgs a) {
  const int32_t lb = logical_block(blockIdx.x, a.xcd_first, a.xcd_offset, a.xcd_remap);
  const int32_t p = lb * BLOCK + threadIdx.x;
  // One wavefront covers one 64-slot chunk: its base loads are scalar.
  const int32_t chunk = __builtin_amdgcn_readfirstlane(p / CHUNK);
  if (p < a.n) ColKern<L, S>::template main_slot<MODE, FORCE>(a, p, chunk);
}

template <class L, class S, int MODE, bool FORCE>
__global__ __launch_bounds__(BLOCK) void boundary_kernel(ColArgs a) {
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < a.n_bslots) ColKern<L, S>::template boundary_slot<MODE, FORCE>(a, a.bslots[i]);
}

// Near-wall stress from the step's input populations fs(t): f(t+1) is the
// pull of fs(t) (near-wall nodes are plain fluid, every link present) and
// f_eq(t) is rebuilt from the moments of fs(t), which BGK conserves:
// rho = sum fs, rho u (1 - alpha/2) = sum c fs - F/2 (Guo forcing).
template <class L, class S, bool FORCE>
__global__ __launch_bounds__(BLOCK) void wss_kernel(ColWssArgs a) {
  using P = Pop<S>;
  using C = typename P::C;
  constexpr int Q = L::Q;
  constexpr bool SHIFT = P::shifted;
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= a.n_nw) return;
  const int32_t p = a.slot[i];
  C F[Q], g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int32_t s = a.src[int64_t(q) * a.n_nw + i];
    F[q] = s >= 0 ? P::ld(a.fs_in, int64_t(q) * a.npad + s) : P::ld(a.fs_in, int64_t(L::opp[q]) * a.npad + p);
    g[q] = P::ld(a.fs_in, int64_t(q) * a.npad + p);
  }
  C s, jx, jy, jz;
  moments<L, C>(g, s, jx, jy, jz);
  const C rho = SHIFT ? C(1) + s : s;
  C den = rho;
  if (FORCE) {
    const C alpha = a.porosity ? C(a.porosity[p]) : C(0);
    jx -= C(0.5) * C(a.force[0]);
    jy -= C(0.5) * C(a.force[1]);
    jz -= C(0.5) * C(a.force[2]);
    den = rho * (C(1) - C(0.5) * alpha);
  }
  const C ux = jx / den, uy = jy / den, uz = jz / den;
  const C usq = ux * ux + uy * uy + uz * uz;
  const double n[3] = {a.normal[3 * i], a.normal[3 * i + 1], a.normal[3 * i + 2]};
  double tau[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int al = 0; al < Q; ++al) {
    const double fne = double(F[al] - geq_of<L, C, SHIFT>(al, s, rho, ux, uy, uz, usq));
    const double cn = L::cx[al] * n[0] + L::cy[al] * n[1] + L::cz[al] * n[2];
    // tau_i += fne (c.n) (c_i - (c.n) n_i)
    tau[0] += fne * cn * (L::cx[al] - cn * n[0]);
    tau[1] += fne * cn * (L::cy[al] - cn * n[1]);
    tau[2] += fne * cn * (L::cz[al] - cn * n[2]);
  }
  const int64_t m = a.n_nw;
  const double t0 = a.wss_C * tau[0], t1 = a.wss_C * tau[1], t2 = a.wss_C * tau[2];
  const double W = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
  a.wss[i] = W;
  a.wss[m + i] += t0;
  a.wss[2 * m + i] += t1;
  a.wss[3 * m + i] += t2;
  a.wss[4 * m + i] += W;
}

__global__ __launch_bounds__(BLOCK) void wss_stats_cols(ColWssArgs a, double* partials) {
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double s = 0.0, mx = 0.0;
  if (i < a.n_nw) {
    const double w = a.wss[i];
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
hipError_t launch_cols(const ColArgs& a, hipStream_t st) {
  if (a.nblocks > 0)
    hipLaunchKernelGGL((columns_kernel<L, S, MODE, FORCE>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

}  // namespace

template <class L, class S, int MODE, bool FORCE>
static hipError_t launch_pair(const ColArgs& a, hipStream_t st, int parts) {
  hipError_t e = hipSuccess;
  if (parts & COLS_MAIN) e = launch_cols<L, S, MODE, FORCE>(a, st);
  if (e != hipSuccess) return e;
  if ((parts & COLS_BOUNDARY) && MODE != KMODE_INIT && a.n_bslots > 0) {
    const unsigned nb = static_cast<unsigned>((a.n_bslots + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL((boundary_kernel<L, S, MODE, FORCE>), dim3(nb), dim3(BLOCK), 0, st, a);
    e = hipGetLastError();
  }
  return e;
}

template <class L, class S>
static hipError_t launch_cols_mode(int kmode, bool force, const ColArgs& a, hipStream_t st, int parts) {
  switch (kmode) {
    case KMODE_STEP:
      return force ? launch_pair<L, S, KMODE_STEP, true>(a, st, parts)
                   : launch_pair<L, S, KMODE_STEP, false>(a, st, parts);
    case KMODE_INIT:
      return force ? launch_pair<L, S, KMODE_INIT, true>(a, st, parts)
                   : launch_pair<L, S, KMODE_INIT, false>(a, st, parts);
    case KMODE_MATERIALIZE:
      return launch_pair<L, S, KMODE_MATERIALIZE, false>(a, st, parts);
  }
  return hipErrorInvalidValue;
}

template <class L>
static hipError_t launch_cols_prec(int precision, int kmode, bool force, const ColArgs& a, hipStream_t st,
                                   int parts) {
  switch (precision) {
    case 0: return launch_cols_mode<L, double>(kmode, force, a, st, parts);
    case 1: return launch_cols_mode<L, float>(kmode, force, a, st, parts);
    case 2: return launch_cols_mode<L, __half>(kmode, force, a, st, parts);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_columns(int Q, int precision, int kmode, bool force, const ColArgs& a, hipStream_t st,
                          int parts) {
  if (Q == 19) return launch_cols_prec<D3Q19>(precision, kmode, force, a, st, parts);
  if (Q == 27) return launch_cols_prec<D3Q27>(precision, kmode, force, a, st, parts);
  return hipErrorInvalidValue;
}

template <class L, class S>
static hipError_t wss_one(bool force, const ColWssArgs& a, hipStream_t st) {
  if (a.n_nw <= 0) return hipSuccess;
  const unsigned nb = static_cast<unsigned>((a.n_nw + BLOCK - 1) / BLOCK);
  if (force) hipLaunchKernelGGL((wss_kernel<L, S, true>), dim3(nb), dim3(BLOCK), 0, st, a);
  else hipLaunchKernelGGL((wss_kernel<L, S, false>), dim3(nb), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_columns_wss(int Q, int precision, bool force, const ColWssArgs& a, hipStream_t st) {
  if (Q == 19) {
    if (precision == 0) return wss_one<D3Q19, double>(force, a, st);
    if (precision == 1) return wss_one<D3Q19, float>(force, a, st);
    if (precision == 2) return wss_one<D3Q19, __half>(force, a, st);
  } else if (Q == 27) {
    if (precision == 0) return wss_one<D3Q27, double>(force, a, st);
    if (precision == 1) return wss_one<D3Q27, float>(force, a, st);
    if (precision == 2) return wss_one<D3Q27, __half>(force, a, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_columns_wss_stats(const ColWssArgs& a, double* partials, double* out2, hipStream_t st) {
  const int32_t nb = (a.n_nw + BLOCK - 1) / BLOCK;
  if (nb == 0) return hipMemsetAsync(out2, 0, 2 * sizeof(double), st);
  hipLaunchKernelGGL(wss_stats_cols, dim3(nb), dim3(BLOCK), 0, st, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_finish(partials, nb, 2, true, out2, st);
}

}  // namespace lbm
