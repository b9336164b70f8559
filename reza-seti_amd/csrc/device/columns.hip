// Column-segment kernel family (see columns.h for the layout).
//
// One launch per reference iteration:
//   leading blocks  inlet/outlet nodes: pull, inlet/outlet rule, collide,
//                   store (a short dependent-load chain; dispatched first so
//                   its latency hides under the rest of the launch);
//   other blocks    every other slot: pull (regular slots from their
//                   segment's 8 lateral deltas, explicit rows from their int32
//                   sources; half-way bounce-back on missing links), collide,
//                   store. Near-wall nodes also accumulate the wall shear
//                   stress (Simulation::ComputeWallShearStress,
//                   Simulation.cpp:767-799) from the pulled populations and
//                   the equilibrium of the collision that produced them.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "colkern.h"

namespace lbm {

namespace {

using namespace dev;
using colk::ColKern;

template <class L, class S, int MODE, bool FORCE, bool NT, bool NTL, bool A32, int KRES>
__device__ inline void columns_body(const ColArgs& a) {
  const int32_t b = blockIdx.x;
  if (b < a.b_blocks) {
    const int32_t i = b * BLOCK + threadIdx.x;
    if (MODE != KMODE_INIT && i < a.n_bslots) ColKern<L, S>::template boundary_slot<MODE, FORCE, NT>(a, a.bslots[i], i);
    return;
  }
  const int32_t lb = main_block_of(a, b);
  if (lb < 0) return;
  const int32_t p = lb * BLOCK + threadIdx.x;
  // One wavefront covers one chunk: its metadata is wave-uniform.
  const int32_t chunk = __builtin_amdgcn_readfirstlane(p / CHUNK);
  const ChunkMeta m = a.meta[chunk];
  if (p < a.n) ColKern<L, S>::template main_slot<MODE, FORCE, NT, NTL, A32, KRES>(a, p, p & (CHUNK - 1), m);
}

template <class L, class S, int MODE, bool FORCE, bool NT, bool NTL, bool A32, int KRES>
__global__ __launch_bounds__(BLOCK) void columns_kernel(ColArgs a) {
  columns_body<L, S, MODE, FORCE, NT, NTL, A32, KRES>(a);
}

// Occupancy-bounded variant for the float-arithmetic precisions: the
// inlet/outlet path needs more registers than the main path, and a
// waves-per-EU bound keeps the main path's occupancy while that rare path
// spills instead. (FP64 keeps the unbounded kernel: its main path spills.)
// 6 waves (80 VGPRs, 3 bytes spilled for D3Q19): 77.7 us per FP32 step over
// 8000 steps of the 2.85M-node bifurcation, against 78.8 us with 7 waves
// (72 VGPRs, 11 bytes spilled) and 91 VGPRs unbounded.
#ifndef LBM_COLUMNS_WAVES_PER_EU
#define LBM_COLUMNS_WAVES_PER_EU 6
#endif
#if LBM_COLUMNS_WAVES_PER_EU > 0
template <class L, class S, int MODE, bool FORCE, bool NT, bool NTL, bool A32, int KRES>
__global__ __attribute__((amdgpu_waves_per_eu(LBM_COLUMNS_WAVES_PER_EU))) __launch_bounds__(BLOCK) void
columns_kernel_occ(ColArgs a) {
  columns_body<L, S, MODE, FORCE, NT, NTL, A32, KRES>(a);
}
#endif

template <class C>
__global__ __launch_bounds__(BLOCK) void wss_flush_kernel(ColArgs a, int32_t count) {
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= a.n_nw) return;
  // Everything here is touched once per flush: nontemporal accesses keep
  // it from displacing a cache-resident state.
  const int64_t m = a.n_nw;
  const colk::Quad<C>* r = static_cast<const colk::Quad<C>*>(a.nw_ring);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  colk::Quad<C> v{};
  for (int32_t k = 0; k < count; ++k) {
    v = colk::load_quad_nt(r + int64_t(k) * m + i);
    acc[0] += double(v.x);
    acc[1] += double(v.y);
    acc[2] += double(v.z);
    acc[3] += double(v.w);
  }
  __builtin_nontemporal_store(double(v.w), a.nw_wss + i);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double* t = a.nw_wss + (1 + c) * m + i;
    __builtin_nontemporal_store(__builtin_nontemporal_load(t) + acc[c], t);
  }
}

__global__ __launch_bounds__(BLOCK) void wss_stats_cols(ColArgs a, double* partials) {
  const int32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double s = 0.0, mx = 0.0;
  if (i < a.n_nw) {
    const double w = a.nw_wss[i];
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

template <class L, class S, int MODE, bool FORCE, bool NT, bool NTL, bool A32 = false, int KRES = 0>
hipError_t launch_one(const ColArgs& a, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>(a.b_blocks + main_grid_blocks(a));
  if (!grid) return hipSuccess;
#if LBM_COLUMNS_WAVES_PER_EU > 0
  if constexpr (!std::is_same<S, double>::value) {
    hipLaunchKernelGGL((columns_kernel_occ<L, S, MODE, FORCE, NT, NTL, A32, KRES>), dim3(grid), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
  }
#endif
  hipLaunchKernelGGL((columns_kernel<L, S, MODE, FORCE, NT, NTL, A32, KRES>), dim3(grid), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

template <class L, class S, bool FORCE, bool A32>
hipError_t launch_step_hint(const ColArgs& a, hipStream_t st) {
  switch (a.cache_hint) {
    case CACHE_NT_STORES:
      if constexpr (A32) {
        if (a.resident_planes == 1) return launch_one<L, S, KMODE_STEP, FORCE, true, false, A32, 1>(a, st);
        if (a.resident_planes == 2) return launch_one<L, S, KMODE_STEP, FORCE, true, false, A32, 2>(a, st);
      }
      return launch_one<L, S, KMODE_STEP, FORCE, true, false, A32>(a, st);
    case CACHE_NT_LOADS:
      if constexpr (A32) {
        if (a.resident_planes == 1) return launch_one<L, S, KMODE_STEP, FORCE, false, true, A32, 1>(a, st);
        if (a.resident_planes == 2) return launch_one<L, S, KMODE_STEP, FORCE, false, true, A32, 2>(a, st);
      }
      return launch_one<L, S, KMODE_STEP, FORCE, false, true, A32>(a, st);
  }
  return launch_one<L, S, KMODE_STEP, FORCE, false, false, A32>(a, st);
}

// 32-bit byte offsets when the whole state buffer spans fewer than 2^32 bytes.
template <class L, class S, bool FORCE>
hipError_t launch_step(const ColArgs& a, hipStream_t st) {
  const bool a32 = !a.wide && uint64_t(L::Q) * uint64_t(a.npad) * sizeof(S) < (uint64_t(1) << 32);
  return a32 ? launch_step_hint<L, S, FORCE, true>(a, st) : launch_step_hint<L, S, FORCE, false>(a, st);
}

// Adjoint recomputation with the moments tape, with the step's cache hint
// (no resident planes).
template <class L, class S, bool FORCE, bool A32>
hipError_t launch_tape_hint(const ColArgs& a, hipStream_t st) {
  switch (a.cache_hint) {
    case CACHE_NT_STORES: return launch_one<L, S, KMODE_TAPE, FORCE, true, false, A32>(a, st);
    case CACHE_NT_LOADS: return launch_one<L, S, KMODE_TAPE, FORCE, false, true, A32>(a, st);
  }
  return launch_one<L, S, KMODE_TAPE, FORCE, false, false, A32>(a, st);
}

template <class L, class S, bool FORCE>
hipError_t launch_tape(const ColArgs& a, hipStream_t st) {
  if (!a.tape || (a.spec_out && !a.spec_flag)) return hipErrorInvalidValue;
  const bool a32 = !a.wide && uint64_t(L::Q) * uint64_t(a.npad) * sizeof(S) < (uint64_t(1) << 32);
  return a32 ? launch_tape_hint<L, S, FORCE, true>(a, st) : launch_tape_hint<L, S, FORCE, false>(a, st);
}

template <class L, class S>
hipError_t launch_mode(int kmode, bool force, const ColArgs& a, hipStream_t st) {
  switch (kmode) {
    case KMODE_STEP:
      return force ? launch_step<L, S, true>(a, st) : launch_step<L, S, false>(a, st);
    case KMODE_TAPE:
      return force ? launch_tape<L, S, true>(a, st) : launch_tape<L, S, false>(a, st);
    case KMODE_INIT:
      return force ? launch_one<L, S, KMODE_INIT, true, false, false>(a, st)
                   : launch_one<L, S, KMODE_INIT, false, false, false>(a, st);
    case KMODE_MATERIALIZE:
      return launch_one<L, S, KMODE_MATERIALIZE, false, false, false>(a, st);
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

hipError_t launch_wss_flush(int precision, const ColArgs& a, int32_t count, hipStream_t st) {
  if (count <= 0 || a.n_nw <= 0) return hipSuccess;
  const unsigned grid = static_cast<unsigned>((a.n_nw + BLOCK - 1) / BLOCK);
  if (precision == 0) hipLaunchKernelGGL(wss_flush_kernel<double>, dim3(grid), dim3(BLOCK), 0, st, a, count);
  else hipLaunchKernelGGL(wss_flush_kernel<float>, dim3(grid), dim3(BLOCK), 0, st, a, count);
  return hipGetLastError();
}

namespace {
// One 16-byte element per thread, nontemporal both ways.
__global__ __launch_bounds__(BLOCK) void stream_copy_kernel(const void* src, void* dst, int64_t n16) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i < n16)
    __builtin_nontemporal_store(__builtin_nontemporal_load(static_cast<const V4*>(src) + i), static_cast<V4*>(dst) + i);
}
}  // namespace

hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes % 16) return hipErrorInvalidValue;
  const int64_t n16 = static_cast<int64_t>(bytes / 16);
  if (n16 == 0) return hipSuccess;
  const int64_t grid = (n16 + BLOCK - 1) / BLOCK;
  if (grid >= (int64_t(1) << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(stream_copy_kernel, dim3(static_cast<unsigned>(grid)), dim3(BLOCK), 0, st, src, dst, n16);
  return hipGetLastError();
}

hipError_t launch_columns_wss_stats(const ColArgs& a, double* partials, double* out2, hipStream_t st) {
  const int32_t nb = (a.n_nw + BLOCK - 1) / BLOCK;
  if (nb == 0) return hipMemsetAsync(out2, 0, 2 * sizeof(double), st);
  hipLaunchKernelGGL(wss_stats_cols, dim3(nb), dim3(BLOCK), 0, st, a, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_finish(partials, nb, 2, true, out2, st);
}

}  // namespace lbm
