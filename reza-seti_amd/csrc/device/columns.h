// Column-offset kernel family: argument block and launchers.
//
// Used for half-way bounce-back walls (the reference has no node that keeps
// stale populations there), with any inlet/outlet rule. Slots are the
// reference node order itself (i outer, k inner), so the node at
// (x', y', z - 1) sits one slot before (x', y', z) whenever both exist. A
// regular node therefore needs only the slots of its 8 lateral column
// neighbours (x - ox, y - oy, z); every pull source is one of them (or the
// node itself) shifted by -c_z. Those 8 slots are stored as 16-bit offsets
// from a per-64-slot base. Nodes that break the rule (missing links,
// periodic wraps), inlet/outlet nodes and near-wall nodes carry an explicit
// int32 row instead; a row also holds the node's boundary data.
//
// One launch per step: the first `b_blocks` blocks handle inlet/outlet nodes
// (one thread each), the rest every other slot, XCD-remapped. Near-wall rows
// accumulate the wall shear stress inside the same launch.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {

constexpr int NCOL = 8;              // lateral column offsets
constexpr int CHUNK = 64;            // slots per base entry (one wavefront)
constexpr uint16_t IRREGULAR = 0xFFFF;
constexpr uint32_t ROW_KIND_MASK = 3u;  // NodeKind
constexpr uint32_t ROW_NEARWALL = 4u;

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
  int32_t b_blocks;      // leading blocks for inlet/outlet slots
  int32_t m_blocks;      // blocks for the other slots
  int32_t xcd_remap;
  int32_t xcd_first[8];
  int32_t xcd_offset[8];
  const uint16_t* idx16;  // NCOL * npad; idx16[0] == IRREGULAR marks an irregular slot,
                          // whose row index is idx16[1] | idx16[2] << 16
  const int32_t* cbase;   // NCOL * nchunk
  int32_t nchunk;
  int32_t n_irr;
  const int32_t* irr_src;    // Q * n_irr pull sources (slot or -1 = bounce back)
  const uint32_t* irr_kind;  // n_irr: NodeKind | ROW_NEARWALL
  const int32_t* irr_bcn;    // n_irr: slot behind an inlet/outlet, or -1
  const int32_t* irr_inlet;  // n_irr: inlet basis index, or -1
  const double* irr_normal;  // 3 * n_irr (near-wall rows)
  double* irr_rho_u;         // 4 * n_irr: rho (or rho - 1), u of the last collision (near-wall rows)
  double* irr_wss;           // 5 * n_irr: |tau|, acc tau x/y/z, acc |tau| (near-wall rows)
  double wss_C;
  int32_t wss_on;
  const int32_t* bslots;     // inlet/outlet slots
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

hipError_t launch_columns(int Q, int precision, int kmode, bool force, const ColArgs& a, hipStream_t st);
// Sum |tau|^2 and max |tau| over near-wall rows (2 doubles into out2);
// partials needs 2 * ceil(n_irr / 256) doubles.
hipError_t launch_columns_wss_stats(const ColArgs& a, double* partials, double* out2, hipStream_t st);

}  // namespace lbm
