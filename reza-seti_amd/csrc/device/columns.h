// Column-offset kernel family: argument blocks and launchers.
//
// Used for half-way bounce-back walls (the reference has no node that keeps
// stale populations there), with any inlet/outlet rule. Slots are the
// reference node order itself (i outer, k inner), so the node at
// (x', y', z - 1) sits one slot before (x', y', z) whenever both exist. A
// regular node therefore needs only the slots of its 8 lateral column
// neighbours (x - ox, y - oy, z); every pull source is one of them (or the
// node itself) shifted by -c_z. Those 8 slots are stored as 16-bit offsets
// from a per-64-slot base. Nodes that break the rule (missing links,
// periodic wraps, inlet/outlet nodes) carry an explicit int32 row.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {

constexpr int NCOL = 8;              // lateral column offsets
constexpr int CHUNK = 64;            // slots per base entry (one wavefront)
constexpr uint16_t IRREGULAR = 0xFFFF;

// Lateral offset index of a direction with (cx, cy), -1 for the own column.
__host__ __device__ constexpr int col_of(int cx, int cy) {
  return (cx == 1 && cy == 0)     ? 0
         : (cx == -1 && cy == 0)  ? 1
         : (cx == 0 && cy == 1)   ? 2
         : (cx == 0 && cy == -1)  ? 3
         : (cx == 1 && cy == 1)   ? 4
         : (cx == 1 && cy == -1)  ? 5
         : (cx == -1 && cy == 1)  ? 6
         : (cx == -1 && cy == -1) ? 7
                                  : -1;
}

struct ColArgs {
  const void* fs_in;
  void* fs_out;
  int64_t npad;          // SoA stride
  int32_t n;             // slots [0, n)
  int32_t nblocks;
  int32_t xcd_remap;
  int32_t xcd_first[8];
  int32_t xcd_offset[8];
  const uint16_t* idx16;  // NCOL * npad; idx16[0] == IRREGULAR marks an irregular slot,
                          // whose row index is idx16[1] | idx16[2] << 16
  const int32_t* cbase;   // NCOL * nchunk
  int32_t nchunk;
  int32_t n_irr;
  const int32_t* irr_src;    // Q * n_irr pull sources (slot or -1 = bounce back)
  const uint32_t* irr_kind;  // n_irr: NodeKind
  const int32_t* irr_bcn;    // n_irr: slot behind an inlet/outlet, or -1
  const int32_t* irr_inlet;  // n_irr: inlet basis index, or -1
  const int32_t* bslots;     // inlet/outlet slots handled by the boundary kernel
  int32_t n_bslots;
  int32_t n_inlet;
  int32_t n_modes;
  const double* inlet_basis;
  double coef[16];
  double inlet_normal[3];
  int32_t inlet_mode;
  int32_t outlet_mode;
  double outlet_rho;
  double omega;
  double force[3];
  const float* porosity;  // per slot, or null
  const double* f_direct; // init input (FP64 unshifted, q * npad + slot), null = w_q
  double* f_out;          // materialize output
};

struct ColWssArgs {
  const void* fs_in;      // input populations of the step just launched
  int64_t npad;
  int32_t n_nw;
  const int32_t* slot;    // n_nw
  const int32_t* src;     // Q * n_nw pull sources
  const double* normal;   // 3 * n_nw
  double* wss;            // 5 * n_nw: |tau|, acc tau x/y/z, acc |tau|
  double wss_C;
  double omega;
  double force[3];
  const float* porosity;
};

// parts: COLS_MAIN (every slot but inlet/outlet nodes), COLS_BOUNDARY
// (inlet/outlet nodes; they read only fs_in and write only their own slots,
// so they may run concurrently with the main launch), or both in order.
constexpr int COLS_MAIN = 1, COLS_BOUNDARY = 2, COLS_ALL = 3;
hipError_t launch_columns(int Q, int precision, int kmode, bool force, const ColArgs& a, hipStream_t st,
                          int parts = COLS_ALL);
hipError_t launch_columns_wss(int Q, int precision, bool force, const ColWssArgs& a, hipStream_t st);
// Sum |tau|^2 and max |tau| (2 doubles into out2); partials needs 2 * ceil(n_nw / 256).
hipError_t launch_columns_wss_stats(const ColWssArgs& a, double* partials, double* out2, hipStream_t st);

}  // namespace lbm
