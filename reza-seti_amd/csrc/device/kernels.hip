// Fused collide-stream kernels for gfx950 (MI355X), reference boundary rules.
//
// One launch = one reference iteration's Streaming + inlet/outlet/wall
// boundary rules + wall shear stress (Simulation.cpp:555-568) followed by the
// next iteration's Collision (Simulation.cpp:77-232). The state between
// launches is the post-collision population f* in SoA layout over a sparse
// slot list; streaming is a pull through an int32 table.
//
// Slots are split in two ranges:
//   special [0, n_special): wall, inlet, outlet and near-wall nodes. They
//     carry side data (previous post-BC f for missing links, boundary
//     neighbour, WSS accumulators) and run the full boundary logic.
//   bulk [n_special_pad, bulk_end): plain fluid nodes with every link
//     present: pull Q values, collide, store Q values.
// Special blocks are dispatched first (they are the heavier ones); bulk
// blocks are remapped so that each XCD walks a contiguous slot range and
// neighbour reads stay in that XCD's L2.
//
// FP64 (precision 0) reproduces the reference arithmetic term by term; the
// library is built with -ffp-contract=off so no FMA is formed. The FP32/FP16
// paths store f - w_q and compute in FP32 with contraction enabled locally.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../lattice.h"
#include "collision.h"
#include "kernels.h"

namespace lbm {

namespace {

using namespace dev;

template <class L, class S>
struct Kern {
  using P = Pop<S>;
  using C = typename P::C;
  static constexpr int Q = L::Q;
  static constexpr bool SHIFT = P::shifted;

  // Population load, nontemporal (NTL) when the state is streamed past the
  // Infinity Cache.
  template <bool NTL = false>
  __device__ static C ld_pop(const void* base, int64_t i) {
    if constexpr (NTL) return P::ld_nt(base, i);
    else return P::ld(base, i);
  }

  // Post-stream populations of slot p (pull), before any boundary rule.
  // Missing links take the slot's previous post-BC value (reference: the
  // target of a push that never happens keeps its value, Simulation.cpp:245-251)
  // or, in half-way bounce-back mode, the slot's own opposite population.
  // OWN: p is the thread's own slot (wave-uniform chunk), so its sources may
  // come from the column segments; otherwise (the node behind an
  // inlet/outlet node) from the table.
  template <bool NTL = false, bool OWN = false>
  __device__ static void gather(const StepArgs& a, int32_t p, C* F, uint32_t& miss) {
    const bool special = p < a.n_special_pad;
    miss = 0;
    int32_t src[Q];
    if constexpr (OWN) {
      slot_sources(a, p, src);
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) src[q] = (q == L::REST) ? p : a.src[int64_t(q) * a.npad + p];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int32_t s = src[q];
      if (s >= 0) {
        C v = ld_pop<NTL>(a.fs_in, int64_t(q) * a.npad + s);
        if (a.zero_check && v == C(0)) {
          // Simulation.cpp:247: a zero f* is not streamed.
          if (special) v = C(a.fprev_in[int64_t(q) * a.n_special_pad + p]);
          else atomicOr(a.err_flags, ERRF_ZERO_SKIP_BULK);
        }
        F[q] = v;
      } else {
        miss |= 1u << q;
        if (a.wall_mode == WALL_BB) F[q] = P::ld(a.fs_in, int64_t(L::opp[q]) * a.npad + p);
        else F[q] = C(a.fprev_in[int64_t(q) * a.n_special_pad + p]);
      }
    }
  }

  __device__ static C inlet_speed(const StepArgs& a, int32_t p) {
    const int32_t i = a.inlet_slot[p];
    double u = 0.0;
    for (int m = 0; m < a.n_modes; ++m) u += a.coef[m] * a.inlet_basis[int64_t(m) * a.n_inlet + i];
    return C(u);
  }

  // Simulation::ComputeWallShearStress for one node (Simulation.cpp:767-799),
  // loops and accumulation order as in the reference; f_eq is rebuilt from the
  // stored density/velocity of the previous collision.
  __device__ static void wss(const StepArgs& a, int32_t p, const C* F) {
    const int64_t sp = a.n_special_pad;
    const double n[3] = {a.normal[p], a.normal[sp + p], a.normal[2 * sp + p]};
    const double r0 = a.rho_u[p], u0 = a.rho_u[sp + p], u1 = a.rho_u[2 * sp + p], u2 = a.rho_u[3 * sp + p];
    double fne[Q];
    if (SHIFT) {
      const C drho = C(r0), rho = C(1) + drho, ux = C(u0), uy = C(u1), uz = C(u2);
      const C usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
      for (int q = 0; q < Q; ++q) fne[q] = double(F[q] - geq_of<L, C, true>(q, drho, rho, ux, uy, uz, usq));
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) fne[q] = double(F[q]) - feq_ref<L>(q, r0, u0, u1, u2);
    }
    double tau[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double sum_alpha = 0.0;
#pragma unroll
      for (int al = 0; al < Q; ++al) {
        const int c[3] = {L::cx[al], L::cy[al], L::cz[al]};
        double sum_k = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) sum_k += c[k] * n[k] * n[i];
        double sum_j = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) sum_j += c[j] * n[j] * (c[i] - sum_k);
        sum_alpha += fne[al] * sum_j;
      }
      tau[i] = a.wss_C * sum_alpha;
    }
    const double W = sqrt(tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]);
    a.wss[p] = W;
    a.wss[sp + p] += tau[0];
    a.wss[2 * sp + p] += tau[1];
    a.wss[3 * sp + p] += tau[2];
    a.wss[4 * sp + p] += W;
  }

  template <int MODE, bool FORCE, bool NT = false>
  __device__ static void collide_store(const StepArgs& a, int32_t p, C* F, bool wall_zero_u, bool nearwall,
                                       bool special) {
    C drho, ux, uy, uz;
    if constexpr (!SHIFT && !FORCE) {
      double rho, dux, duy, duz;
      collide_ref<L>(F, wall_zero_u, a.omega, rho, dux, duy, duz);
      drho = C(rho);
      ux = C(dux);
      uy = C(duy);
      uz = C(duz);
    } else {
      const C frc[3] = {C(a.force[0]), C(a.force[1]), C(a.force[2])};
      const C alpha = a.porosity ? C(a.porosity[p]) : C(0);
      collide_ext<L, C, SHIFT, FORCE>(F, C(a.omega), frc, alpha, wall_zero_u, drho, ux, uy, uz);
    }
    if (special && nearwall) {
      const int64_t sp = a.n_special_pad;
      a.rho_u[p] = double(drho);
      a.rho_u[sp + p] = double(ux);
      a.rho_u[2 * sp + p] = double(uy);
      a.rho_u[3 * sp + p] = double(uz);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if constexpr (NT) P::st_nt(a.fs_out, int64_t(q) * a.npad + p, F[q]);
      else P::st(a.fs_out, int64_t(q) * a.npad + p, F[q]);
    }
  }

  __device__ static void load_direct(const StepArgs& a, int32_t p, C* F) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double f = a.f_direct ? a.f_direct[int64_t(q) * a.npad + p] : L::w[q];
      F[q] = SHIFT ? C(f - L::w[q]) : C(f);
    }
  }

  __device__ static void write_direct(const StepArgs& a, int32_t p, const C* F) {
#pragma unroll
    for (int q = 0; q < Q; ++q) a.f_out[int64_t(q) * a.npad + p] = SHIFT ? double(F[q]) + L::w[q] : double(F[q]);
  }

  // NT (a state larger than the Infinity Cache): the populations are loaded
  // and stored nontemporally and so are the special slots' previous values,
  // so that the int32 pull table, read whole every step, stays in the
  // Infinity Cache (2.85M-node bifurcation, FP64: 344.5 -> 335.1 us per step
  // against nontemporal population stores alone).
  template <int MODE, bool FORCE, bool NT = false>
  __device__ static void special_node(const StepArgs& a, int32_t p) {
    const uint32_t fl = a.sflags[p];
    if (fl & SFLAG_PAD) return;
    const int kind = int(fl & SFLAG_KIND_MASK);
    const bool nearwall = (fl & SFLAG_NEARWALL) != 0;
    C F[Q];
    if (MODE == KMODE_INIT) {
      load_direct(a, p, F);
    } else {
      uint32_t miss;
      gather<NT, true>(a, p, F, miss);
      if (kind == 2 /*inlet*/) {
        if (a.inlet_mode == INLET_ZOUHE) {
          if constexpr (Q == 19) zouhe_ref<C, SHIFT>(F, inlet_speed(a, p), a.inlet_normal);
        } else if (a.inlet_mode == INLET_NEE) {
          const int32_t u = a.bc_nbr[p];
          if (u >= 0) {
            C G[Q];
            uint32_t gm;
            gather(a, u, G, gm);
            const C U = inlet_speed(a, p);
            nee<L, C, SHIFT>(F, G, C(-1), -C(a.inlet_normal[0]) * U, -C(a.inlet_normal[1]) * U,
                             -C(a.inlet_normal[2]) * U, false);
          }
        }
      } else if (kind == 3 /*outlet*/) {
        const int32_t u = a.bc_nbr[p];
        if (u >= 0) {
          if (a.outlet_mode == OUTLET_REF) {
            // Simulation::ApplyOutletBc: copy every population of the node
            // behind the outlet after its own stream and inlet rule.
            C G[Q];
            uint32_t gm;
            gather(a, u, G, gm);
            if (u < a.n_special_pad && int(a.sflags[u] & SFLAG_KIND_MASK) == 2 && a.inlet_mode == INLET_ZOUHE) {
              if constexpr (Q == 19) zouhe_ref<C, SHIFT>(G, inlet_speed(a, u), a.inlet_normal);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) F[q] = G[q];
          } else if (a.outlet_mode == OUTLET_NEE) {
            C G[Q];
            uint32_t gm;
            gather(a, u, G, gm);
            nee<L, C, SHIFT>(F, G, C(a.outlet_rho), C(0), C(0), C(0), true);
          }
        }
      } else if (kind == 1 /*wall*/ && a.wall_mode == WALL_REF) {
        // Simulation::ApplyWallBc: for q in order, if the push neighbour
        // x + c_q is absent (pull source of opp(q) missing), f[opp] = f[q].
#pragma unroll
        for (int q = 0; q < Q; ++q)
          if (miss & (1u << L::opp[q])) F[L::opp[q]] = F[q];
      }
      if (MODE == KMODE_STEP && nearwall && a.wss_on) wss(a, p, F);
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    // Reference post-BC f of boundary nodes, for links that stream nothing.
    if (a.fprev_out) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if constexpr (NT) __builtin_nontemporal_store(double(F[q]), a.fprev_out + int64_t(q) * a.n_special_pad + p);
        else a.fprev_out[int64_t(q) * a.n_special_pad + p] = double(F[q]);
      }
    }
    const bool zero_u = (kind == 1) && (a.wall_mode == WALL_REF);
    collide_store<MODE, FORCE, NT>(a, p, F, zero_u, nearwall, true);
  }

  // Pull sources of slot p (wave-uniform chunk): p + delta(o) - c_z from its
  // column segment (p - c_z in its own column), or its row of the pull table
  // (-1 for a missing link) where the chunk marks the slot irregular (the host
  // checked every regular slot's decode against the table). The branch keeps
  // the table loads off the regular lanes.
  __device__ static void slot_sources(const StepArgs& a, int32_t p, int32_t* src) {
    bool row = true;
    int4 lo = make_int4(0, 0, 0, 0), hi = lo;
    if (a.bmeta) {
      const int32_t c = __builtin_amdgcn_readfirstlane(p / CHUNK);
      const ChunkMeta m = a.bmeta[c];
      const int l = p & (CHUNK - 1);
      row = (m.irr_mask >> l) & 1u;
      const int4* d4 = reinterpret_cast<const int4*>(a.bdelta) + 2 * int64_t(segment_of(m, l));
      lo = d4[0];
      hi = d4[1];
    }
    if (row) {
#pragma unroll
      for (int q = 0; q < Q; ++q) src[q] = (q == L::REST) ? p : a.src[int64_t(q) * a.npad + p];
    } else {
      const int32_t d[NCOL] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int o = col_of(L::cx[q], L::cy[q]);
        src[q] = (o < 0 ? p : p + d[o]) - L::cz[q];
      }
    }
  }

  template <int MODE, bool FORCE, bool NT = false>
  __device__ static void bulk_node(const StepArgs& a, int32_t p) {
    C F[Q];
    if (MODE == KMODE_INIT) {
      load_direct(a, p, F);
    } else {
      int32_t src[Q];
      slot_sources(a, p, src);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const C v = ld_pop<NT>(a.fs_in, int64_t(q) * a.npad + src[q]);
        if (!SHIFT && a.zero_check && v == C(0)) atomicOr(a.err_flags, ERRF_ZERO_SKIP_BULK);
        F[q] = v;
      }
    }
    if (MODE == KMODE_MATERIALIZE) {
      write_direct(a, p, F);
      return;
    }
    collide_store<MODE, FORCE, NT>(a, p, F, false, false, false);
  }
};

__device__ inline int32_t bulk_chunk(const StepArgs& a, int32_t b) {
  if (!a.xcd_remap) return b - a.blocks_special;
  const int x = b & 7;
  return a.xcd_offset[x] + (b - a.xcd_first[x]) / 8;
}

template <class L, class S, int MODE, bool FORCE>
__global__ __launch_bounds__(BLOCK) void lbm_kernel(StepArgs a) {
  const int32_t b = blockIdx.x;
  if (b < a.blocks_special) {
    const int32_t p = b * BLOCK + threadIdx.x;
    if (p < a.n_special_pad) Kern<L, S>::template special_node<MODE, FORCE>(a, p);
    return;
  }
  const int32_t p = a.n_special_pad + bulk_chunk(a, b) * BLOCK + threadIdx.x;
  if (p < a.bulk_end) Kern<L, S>::template bulk_node<MODE, FORCE>(a, p);
}

// ------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-level sum of NV doubles; thread 0 writes partials[blockIdx * NV + i].
template <int NV>
__device__ inline void block_sum_store(double* v, double* partials) {
  __shared__ double red[BLOCK / 64][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wv][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < BLOCK / 64; ++w) s += red[w][i];
      partials[int64_t(blockIdx.x) * NV + i] = s;
    }
}

template <class L, class S>
__global__ __launch_bounds__(BLOCK) void moments_kernel(StepArgs a, double* partials) {
  using P = Pop<S>;
  const int32_t p = blockIdx.x * BLOCK + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const bool valid = (p < a.n_special_pad) ? (p < a.n_special && !(a.sflags[p] & SFLAG_PAD))
                                           : (p < a.bulk_end);
  if (valid) {
    double s = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      double f = double(P::ld(a.fs_in, int64_t(q) * a.npad + p));
      if (P::shifted) f += L::w[q];
      s += f;
      jx += L::cx[q] * f;
      jy += L::cy[q] * f;
      jz += L::cz[q] * f;
    }
    v[0] = s;
    v[1] = jx;
    v[2] = jy;
    v[3] = jz;
  }
  block_sum_store<4>(v, partials);
}

__global__ __launch_bounds__(BLOCK) void wss_stats_kernel(StepArgs a, double* partials) {
  const int32_t p = blockIdx.x * BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (p < a.n_special && (a.sflags[p] & SFLAG_NEARWALL) && !(a.sflags[p] & SFLAG_PAD)) {
    const double w = a.wss[p];
    v[0] = w * w;
    v[1] = w;
  }
  __shared__ double red[BLOCK / 64][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v[0] = wave_sum(v[0]);
  v[1] = wave_max(v[1]);
  if (lane == 0) {
    red[wv][0] = v[0];
    red[wv][1] = v[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0, m = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      s += red[w][0];
      m = fmax(m, red[w][1]);
    }
    partials[2 * blockIdx.x] = s;
    partials[2 * blockIdx.x + 1] = m;
  }
}

// Fixed-order final pass (deterministic): one block folds nblk partials.
template <int NV, bool MAXLAST>
__global__ __launch_bounds__(BLOCK) void finish_kernel(const double* partials, int32_t nblk, double* out) {
  __shared__ double red[BLOCK][NV];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  for (int32_t b = threadIdx.x; b < nblk; b += BLOCK)
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const double x = partials[int64_t(b) * NV + i];
      acc[i] = (MAXLAST && i == NV - 1) ? fmax(acc[i], x) : acc[i] + x;
    }
#pragma unroll
  for (int i = 0; i < NV; ++i) red[threadIdx.x][i] = acc[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int t = 0; t < BLOCK; ++t) s = (MAXLAST && i == NV - 1) ? fmax(s, red[t][i]) : s + red[t][i];
      out[i] = s;
    }
  }
}

// The two launches of the FP64 step (launch_split); together they write
// what one lbm_kernel launch writes. The special launch is bounded to 4 waves
// per SIMD (128 VGPRs instead of 168, some spills on the heavier paths):
// 355 -> 345 us per step on the 2.85M-node bifurcation.
template <class L, class S, int MODE, bool FORCE, bool NT>
__global__ __attribute__((amdgpu_waves_per_eu(4))) __launch_bounds__(BLOCK) void lbm_special_kernel(StepArgs a) {
  const int32_t p = blockIdx.x * BLOCK + threadIdx.x;
  if (p < a.n_special_pad) Kern<L, S>::template special_node<MODE, FORCE, NT>(a, p);
}

// The bulk launch walks runs of 32 blocks per XCD (8192 slots; 274.3 us per
// step on the 2.85M-node bifurcation against 276.1 us with one contiguous
// eighth of the bulk blocks per XCD and 275.7 us with runs of 64).
constexpr int32_t BULK_XCD_RUN = 32;

template <class L, class S, int MODE, bool FORCE, bool NT>
__global__ __launch_bounds__(BLOCK) void lbm_bulk_kernel(StepArgs a) {
  // Block b runs on XCD b mod 8; XCD x takes the runs x, x + 8, x + 16, ...
  // of BULK_XCD_RUN blocks, and the blocks that round the grid up to whole
  // rounds of 8 runs do nothing.
  const int32_t b = blockIdx.x;
  const int32_t k = b >> 3;
  const int32_t chunk = (k / BULK_XCD_RUN) * 8 * BULK_XCD_RUN + (b & 7) * BULK_XCD_RUN + k % BULK_XCD_RUN;
  if (chunk >= a.blocks_bulk) return;
  const int32_t p = a.n_special_pad + chunk * BLOCK + threadIdx.x;
  if (p < a.bulk_end) Kern<L, S>::template bulk_node<MODE, FORCE, NT>(a, p);
}

#ifndef LBM_REF_SPLIT
#define LBM_REF_SPLIT 1
#endif

// The FP64 step as two launches: the special slots (boundary rules, stress,
// previous values; many registers) and the bulk slots (pull, collide, store;
// few registers, so many more wavefronts in flight). (One launch per special
// group, each with only its own rule's registers, took 419 us per step on the
// 2.85M-node bifurcation against 355 us: the inlet/outlet group alone is a
// 51 us chain of dependent loads, which the combined launch hides.)
template <class L, class S, int MODE, bool FORCE, bool NT>
hipError_t launch_split(const StepArgs& a, hipStream_t st) {
  if (a.blocks_special)
    hipLaunchKernelGGL((lbm_special_kernel<L, S, MODE, FORCE, NT>), dim3(a.blocks_special), dim3(BLOCK), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.blocks_bulk) return e;
  const int32_t round = 8 * BULK_XCD_RUN;
  hipLaunchKernelGGL((lbm_bulk_kernel<L, S, MODE, FORCE, NT>), dim3((a.blocks_bulk + round - 1) / round * round),
                     dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

template <class L, class S, int MODE, bool FORCE>
hipError_t launch_one(const StepArgs& a, hipStream_t st) {
  if constexpr (LBM_REF_SPLIT && MODE == KMODE_STEP && std::is_same<S, double>::value) {
    return a.nt_stores ? launch_split<L, S, MODE, FORCE, true>(a, st) : launch_split<L, S, MODE, FORCE, false>(a, st);
  }
  const dim3 grid(a.blocks_special + a.blocks_bulk);
  if (grid.x == 0) return hipSuccess;
  hipLaunchKernelGGL((lbm_kernel<L, S, MODE, FORCE>), grid, dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

template <class L, class S>
hipError_t launch_mode(int kmode, bool force, const StepArgs& a, hipStream_t st) {
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
hipError_t launch_prec(int precision, int kmode, bool force, const StepArgs& a, hipStream_t st) {
  switch (precision) {
    case 0: return launch_mode<L, double>(kmode, force, a, st);
    case 1: return launch_mode<L, float>(kmode, force, a, st);
    case 2: return launch_mode<L, __half>(kmode, force, a, st);
  }
  return hipErrorInvalidValue;
}

template <class L, class S>
hipError_t moments_one(const StepArgs& a, double* partials, double* out4, hipStream_t st) {
  const int32_t nblk = static_cast<int32_t>((a.bulk_end + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL((moments_kernel<L, S>), dim3(nblk), dim3(BLOCK), 0, st, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((finish_kernel<4, false>), dim3(1), dim3(BLOCK), 0, st, partials, nblk, out4);
  return hipGetLastError();
}

// Density and velocity (FP64 sums) of the state buffer's populations at
// the given slots: 4 values per entry.
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void moments_at_kernel(const void* fs, int64_t npad, const int32_t* slots,
                                                           int64_t count, double* out) {
  using P = Pop<S>;
  const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i >= count) return;
  const int64_t p = slots[i];
  double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    double f = double(P::ld(fs, int64_t(q) * npad + p));
    if (P::shifted) f += L::w[q];
    r += f;
    jx += L::cx[q] * f;
    jy += L::cy[q] * f;
    jz += L::cz[q] * f;
  }
  out[4 * i] = r;
  out[4 * i + 1] = jx / r;
  out[4 * i + 2] = jy / r;
  out[4 * i + 3] = jz / r;
}

template <class L>
hipError_t moments_at_prec(int precision, const void* fs, int64_t npad, const int32_t* slots, int64_t count,
                           double* out, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((count + BLOCK - 1) / BLOCK);
  if (precision == 0) hipLaunchKernelGGL((moments_at_kernel<L, double>), dim3(grid), dim3(BLOCK), 0, st, fs, npad, slots, count, out);
  else if (precision == 1) hipLaunchKernelGGL((moments_at_kernel<L, float>), dim3(grid), dim3(BLOCK), 0, st, fs, npad, slots, count, out);
  else if (precision == 2) hipLaunchKernelGGL((moments_at_kernel<L, __half>), dim3(grid), dim3(BLOCK), 0, st, fs, npad, slots, count, out);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

struct Normal3 {
  double n[3];
};

// Boundary::ApplyBc on independent population vectors (D3Q19, FP64).
__global__ __launch_bounds__(BLOCK) void zouhe_batch_kernel(double* f, const double* vel, int64_t count, Normal3 nrm) {
  const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i >= count) return;
  double m[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) m[q] = f[i * 19 + q];
  zouhe_ref<double, false>(m, vel[i], nrm.n);
#pragma unroll
  for (int q = 0; q < 19; ++q) f[i * 19 + q] = m[q];
}

}  // namespace

hipError_t launch_moments_at(int Q, int precision, const void* fs, int64_t npad, const int32_t* slots, int64_t count,
                             double* out, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (Q == 19) return moments_at_prec<D3Q19>(precision, fs, npad, slots, count, out, st);
  if (Q == 27) return moments_at_prec<D3Q27>(precision, fs, npad, slots, count, out, st);
  return hipErrorInvalidValue;
}

hipError_t launch_zouhe_batch(double* f, const double* vel, int64_t count, const double normal[3], hipStream_t st) {
  if (count <= 0) return hipSuccess;
  Normal3 n{{normal[0], normal[1], normal[2]}};
  const unsigned grid = static_cast<unsigned>((count + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(zouhe_batch_kernel, dim3(grid), dim3(BLOCK), 0, st, f, vel, count, n);
  return hipGetLastError();
}

hipError_t launch_finish(const double* partials, int32_t nblk, int nv, bool maxlast, double* out, hipStream_t st) {
  if (nv == 4 && !maxlast) hipLaunchKernelGGL((finish_kernel<4, false>), dim3(1), dim3(BLOCK), 0, st, partials, nblk, out);
  else if (nv == 2 && maxlast) hipLaunchKernelGGL((finish_kernel<2, true>), dim3(1), dim3(BLOCK), 0, st, partials, nblk, out);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_lbm(int Q, int precision, int kmode, bool force, const StepArgs& a, hipStream_t st) {
  if (Q == 19) return launch_prec<D3Q19>(precision, kmode, force, a, st);
  if (Q == 27) return launch_prec<D3Q27>(precision, kmode, force, a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_moments(int Q, int precision, const StepArgs& a, double* partials, double* out4,
                          hipStream_t st) {
  if (Q == 19) {
    if (precision == 0) return moments_one<D3Q19, double>(a, partials, out4, st);
    if (precision == 1) return moments_one<D3Q19, float>(a, partials, out4, st);
    if (precision == 2) return moments_one<D3Q19, __half>(a, partials, out4, st);
  } else if (Q == 27) {
    if (precision == 0) return moments_one<D3Q27, double>(a, partials, out4, st);
    if (precision == 1) return moments_one<D3Q27, float>(a, partials, out4, st);
    if (precision == 2) return moments_one<D3Q27, __half>(a, partials, out4, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_wss_stats(const StepArgs& a, double* partials, double* out2, hipStream_t st) {
  const int32_t nblk = (a.n_special_pad + BLOCK - 1) / BLOCK;
  if (nblk == 0) return hipMemsetAsync(out2, 0, 2 * sizeof(double), st);
  hipLaunchKernelGGL(wss_stats_kernel, dim3(nblk), dim3(BLOCK), 0, st, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((finish_kernel<2, true>), dim3(1), dim3(BLOCK), 0, st, partials, nblk, out2);
  return hipGetLastError();
}

}  // namespace lbm
