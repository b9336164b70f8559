// Kernel argument block and launch entry points of the fused collide-stream
// kernels. Shared by kernels.hip (device code) and runtime.hip (C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "columns.h"

namespace lbm {

// Per-slot flags of the special-node range.
constexpr uint32_t SFLAG_KIND_MASK = 3u;  // NodeKind
constexpr uint32_t SFLAG_NEARWALL = 4u;
constexpr uint32_t SFLAG_PAD = 8u;        // unused slot

// Kernel variants.
constexpr int KMODE_STEP = 0;         // stream + BC + WSS + collide
constexpr int KMODE_INIT = 1;         // collide f given directly (set / create)
constexpr int KMODE_MATERIALIZE = 2;  // stream + BC only, write post-BC f
constexpr int KMODE_TAPE = 3;         // KMODE_STEP without stress, plus the moments of the pulled
                                       // populations (adjoint recomputation; column family only)

constexpr int BLOCK = 256;
constexpr int MAX_MODES = 16;

// Device error flags.
constexpr uint32_t ERRF_ZERO_SKIP_BULK = 1u;  // f*==0 skip hit a node without kept f

struct StepArgs {
  // Populations: SoA, element (q, slot) at q * npad + slot, storage type per
  // precision (double unshifted, float/half shifted by w_q).
  const void* fs_in;
  void* fs_out;
  const int32_t* src;  // pull table q * npad + slot: slot of x - c_q, or -1
  // Slots as column segments: one ChunkMeta per 64 slots (seg_first,
  // seg_mask; irr_mask: the slot reads its row of `src`) and NCOL deltas per
  // segment; null: every slot reads `src`.
  const ChunkMeta* bmeta;
  const int32_t* bdelta;
  int64_t npad;
  // Slot ranges: special nodes [0, n_special), padding up to n_special_pad,
  // bulk nodes [n_special_pad, bulk_end).
  int32_t n_special;
  int32_t n_special_pad;
  int32_t bulk_end;
  int32_t blocks_special;
  int32_t blocks_bulk;
  int32_t xcd_remap;
  int32_t xcd_first[8];   // first bulk block index dispatched to XCD x
  int32_t xcd_offset[8];  // bulk chunk offset of XCD x
  // Special-node side arrays, index = slot < n_special_pad.
  const uint32_t* sflags;
  const int32_t* bc_nbr;      // slot behind an inlet/outlet node, or -1
  const int32_t* inlet_slot;  // index into the inlet basis, or -1
  int32_t n_inlet;
  int32_t n_modes;
  const double* inlet_basis;  // n_modes * n_inlet
  double coef[MAX_MODES];
  double inlet_normal[3];
  const double* fprev_in;     // Q * n_special_pad: post-BC f of the last step
  double* fprev_out;
  const double* normal;       // 3 * n_special_pad
  double* rho_u;              // 4 * n_special_pad: rho (or rho-1), u of the last collision
  double* wss;                // 5 * n_special_pad: |tau|, acc tau x/y/z, acc |tau|
  double wss_C;
  int32_t wss_on;
  // Collision.
  double omega;
  double force[3];
  const float* porosity;      // npad (Brinkman), or null
  // Boundary modes (LBM_WALL_*, LBM_INLET_*, LBM_OUTLET_* of the C ABI).
  int32_t wall_mode;
  int32_t inlet_mode;
  int32_t outlet_mode;
  double outlet_rho;
  int32_t zero_check;
  int32_t nt_stores;  // 1: the FP64 step stores populations nontemporally (state beyond the Infinity Cache)
  uint32_t* err_flags;
  // KMODE_INIT input / KMODE_MATERIALIZE output: FP64 unshifted f, q * npad + slot.
  const double* f_direct;
  double* f_out;
};

// precision: 0 = f64, 1 = f32 shifted, 2 = f16 shifted. force: body force or
// Brinkman field active.
hipError_t launch_lbm(int Q, int precision, int kmode, bool force, const StepArgs& a, hipStream_t st);

// Mass/momentum over the state in a.fs_in, written as 4 FP64 values to out4
// (device pointer). partials needs 4 * blocks doubles (blocks = n blocks of a).
hipError_t launch_moments(int Q, int precision, const StepArgs& a, double* partials, double* out4,
                          hipStream_t st);

// Sum of |tau|^2 and max |tau| over near-wall slots, 2 FP64 values to out2.
hipError_t launch_wss_stats(const StepArgs& a, double* partials, double* out2, hipStream_t st);

// Deterministic fold of nblk partial rows of nv doubles (nv = 4: sums;
// nv = 2 with maxlast: sum, max) into out.
// Density and velocity (rho, jx/rho, jy/rho, jz/rho) of a state buffer at
// `count` slots, 4 FP64 values per slot to out.
hipError_t launch_moments_at(int Q, int precision, const void* fs, int64_t npad, const int32_t* slots, int64_t count,
                             double* out, hipStream_t st);

// Boundary::ApplyBc (Boundary.cpp:111-289) on `count` D3Q19 FP64 vectors
// f[i * 19 + q], in place, inlet speed vel[i], one normal.
hipError_t launch_zouhe_batch(double* f, const double* vel, int64_t count, const double normal[3], hipStream_t st);

hipError_t launch_finish(const double* partials, int32_t nblk, int nv, bool maxlast, double* out, hipStream_t st);

}  // namespace lbm
