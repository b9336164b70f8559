// Column-segment kernel family: argument block and launchers.
//
// Used for half-way bounce-back walls (the reference has no node that keeps
// stale populations there), with any inlet/outlet rule. Slots are the
// reference node order itself (i outer, k inner), so a run of nodes with the
// same (i, j) and consecutive k -- a column segment -- occupies consecutive
// slots, and the node at (x, y, z - 1) sits one slot before (x, y, z). For
// every segment and each of the 8 lateral offsets (ox, oy), the slot of the
// node at (x - ox, y - oy, z) is the node's own slot plus one per-segment
// delta. A pull source is then p + delta(o) - c_z (or p - c_z in the own
// column), and a node needs no per-node index data at all.
//
// Per 64-slot chunk (one wavefront) a ChunkMeta holds bit masks of the slots
// that start a segment, that miss a link (bounce-back mask in `bbm`), that
// carry an explicit pull row (periodic wraps, segment conflicts, inlet/outlet
// nodes) and that are near-wall (stress data); a lane finds its entry in each
// compact array by a population count over the lower lanes.
//
// One launch per step: the first `b_blocks` blocks handle inlet/outlet nodes
// (one thread each), the rest every other slot, XCD-remapped (each XCD walks
// every eighth run of xcd_run blocks). Near-wall nodes accumulate the wall
// shear stress inside the same launch.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {

constexpr int NCOL = 8;              // lateral column offsets
constexpr int CHUNK = 64;            // slots per ChunkMeta (one wavefront)
constexpr uint32_t ROW_KIND_MASK = 3u;  // NodeKind of an explicit row

// Cache hints of a step's population traffic (the Infinity Cache holds one
// FP32 state of a few million nodes but not two):
constexpr int32_t CACHE_PLAIN = 0;     // both buffers fit: cache everything
constexpr int32_t CACHE_NT_STORES = 1; // stream the output past the cache
constexpr int32_t CACHE_NT_LOADS = 2;  // stream the input, keep the output cached

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

struct alignas(16) ChunkMeta {
  uint64_t seg_mask;   // bit l (l >= 1): slot l starts a new segment
  uint64_t bb_mask;    // regular slots with a missing link
  uint64_t irr_mask;   // slots with an explicit pull row
  uint64_t nw_mask;    // near-wall slots
  int32_t seg_first;   // segment of slot 0 of the chunk
  int32_t bb_base;     // bbm index of the chunk's first bounce-back slot
  int32_t irr_base;    // row of the chunk's first irregular slot
  int32_t nw_base;     // near-wall index of the chunk's first near-wall slot
};
static_assert(sizeof(ChunkMeta) == 48, "ChunkMeta layout");

__host__ __device__ inline int32_t popcount64(uint64_t m) { return __builtin_popcountll(m); }

// Number of set bits of m below lane l (0 <= l < 64).
__host__ __device__ inline int32_t rank_below(uint64_t m, int l) {
  return l ? popcount64(m & (~uint64_t(0) >> (64 - l))) : 0;
}

// Segment of lane l in a chunk: the chunk's first segment plus the segment
// starts at lanes 1..l.
__host__ __device__ inline int32_t segment_of(const ChunkMeta& m, int l) {
  return m.seg_first + popcount64(m.seg_mask & (~uint64_t(1)) & (~uint64_t(0) >> (63 - l)));
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
  int32_t xcd_run;       // > 0: XCD x takes runs x, x + 8, x + 16, ... of xcd_run main blocks
  const ChunkMeta* meta;     // npad / CHUNK
  const int32_t* seg_delta;  // NCOL per segment (32-byte aligned rows)
  const uint32_t* bbm;       // missing-link masks (bit q), bb_mask order
  int32_t n_irr;
  const int32_t* irr_src;    // Q * n_irr pull sources (slot or -1 = bounce back)
  const uint32_t* irr_kind;  // n_irr: NodeKind
  const int32_t* irr_bcn;    // n_irr: slot behind an inlet/outlet, or -1
  const int32_t* irr_inlet;  // n_irr: inlet basis index, or -1
  int32_t n_nw;
  const double* nw_normal;   // 3 * n_nw (SoA), near-wall slot order
  const void* nw_normal_c;   // n_nw x (nx, ny, nz, 0) in the arithmetic precision
  double* nw_wss;            // 5 * n_nw: |tau|, acc tau x/y/z, acc |tau| (FP64)
  void* nw_ring;             // ring slots of n_nw x (tau x, y, z, |tau|), arithmetic precision
  int32_t ring_slot;         // slot this step writes
  void* nw_rho_u;            // n_nw x (density (deviation), ux, uy, uz) of the last collision, arithmetic precision
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
  int32_t cache_hint;     // CACHE_PLAIN, CACHE_NT_STORES or CACHE_NT_LOADS
  int32_t resident_planes;  // 0, 1 or 2: planes [0, k) of both state buffers stay cached (plain accesses)
  const float* porosity;  // per slot, or null
  const double* f_direct; // init input (FP64 unshifted, q * npad + slot), null = w_q
  double* f_out;          // materialize output
  // Objective accumulation (adjoint forward pass), all null otherwise.
  double* obj_acc;        // n_nw: sum over steps of obj_w |tau|^2
  const double* obj_w;    // n_nw: wall-segment weight (0 = outside)
  double* obj_rho;        // n_bslots: sum over steps of the post-rule density
  // KMODE_TAPE output: density (deviation) sum s and momentum jx, jy, jz of
  // every main slot's pulled populations, 4 planes of npad in the
  // arithmetic precision (inlet/outlet slots are not written).
  void* tape;
  // KMODE_TAPE: post-collision populations of the slots flagged in spec_flag
  // (npad bytes) also go to spec_out (a state buffer written only there),
  // or spec_out is null.
  void* spec_out;
  const uint8_t* spec_flag;
  int32_t wide;  // 1: 64-bit element offsets whatever the state size (the path of states of 4 GiB or more)
};

// Blocks of the main part of a launch: m_blocks, rounded up to whole rounds
// of 8 runs when the XCD runs apply.
__host__ __device__ inline int64_t main_grid_blocks(const ColArgs& a) {
  if (!a.xcd_remap || a.xcd_run <= 0) return a.m_blocks;
  const int64_t round = 8 * int64_t(a.xcd_run);
  return (a.m_blocks + round - 1) / round * round;
}

// Main block that grid block b (>= b_blocks) takes, or -1 for a block of the
// rounding. The hardware hands blocks to the 8 XCDs in turn (block b to XCD
// b mod 8); k counts XCD x's blocks in dispatch order.
__device__ inline int32_t main_block_of(const ColArgs& a, int32_t b) {
  if (!a.xcd_remap) return b - a.b_blocks;
  const int x = b & 7;
  const int32_t k = (b - a.xcd_first[x]) / 8;
  if (a.xcd_run <= 0) return a.xcd_offset[x] + k;
  const int32_t lb = (k / a.xcd_run) * 8 * a.xcd_run + x * a.xcd_run + k % a.xcd_run;
  return lb < a.m_blocks ? lb : -1;
}

hipError_t launch_columns(int Q, int precision, int kmode, bool force, const ColArgs& a, hipStream_t st);
// Device copy of `bytes` (a multiple of 16) with nontemporal loads and stores:
// state copies that are not read again soon (checkpoints) pass the caches
// by, instead of taking the Infinity Cache room of a cache-resident state.
hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
// Adds ring slots 0 .. count-1 (in order) to the FP64 accumulators and sets
// the instantaneous |tau| from slot count-1.
hipError_t launch_wss_flush(int precision, const ColArgs& a, int32_t count, hipStream_t st);
// Sum |tau|^2 and max |tau| over near-wall nodes (2 doubles into out2);
// partials needs 2 * ceil(n_nw / 256) doubles.
hipError_t launch_columns_wss_stats(const ColArgs& a, double* partials, double* out2, hipStream_t st);

}  // namespace lbm
