// Device half of the C ABI: solver handle, slot layout, launches.
//
// Replaces the state and time loop of Simulation (Simulation.cpp:44-58,
// 538-618) for a single rank. The MPI halo exchange (Simulation::Communicate,
// Simulation.cpp:270-516) has nothing to do on one GPU: every node lives in
// one slot list and streaming reads neighbours directly.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../capi_internal.h"
#include "../host/checkpoint.h"
#include "../host/output.h"
#include "../host/parallel.h"
#include "../lattice.h"
#include "adjoint.h"
#include "columns.h"
#include "kernels.h"

using namespace lbm;

// Workspace of lbm_reduce_noalloc: the partial sums of its reduction kernels
// and their folded results, in a static device array instead of memory the
// solver allocates. Each GPU holds its own copy from the moment the library's
// code object is loaded on it until the process ends. In host code `scratch`
// only names the symbol: kernels get the device address that
// hipGetSymbolAddress returns for the current device (scratch_on).
constexpr size_t kScratchDoubles = size_t(1) << 23;  // 64 MiB per GPU
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `dev`, which must be the current device;
// looked up once per device.
int scratch_on(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(LBM_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}

}  // namespace

namespace {

float half_to_float(uint16_t h) {
  const uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0) v = std::ldexp(float(m), -24);
  else if (e == 31) v = m ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
  else v = std::ldexp(float(m | 0x400u), int(e) - 25);
  return s ? -v : v;
}

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail(LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));               \
  } while (0)

}  // namespace

struct lbm_solver {
  lbm_params prm{};
  int Q = 19;
  int prec = 0;
  size_t esize = 8;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  int64_t N = 0, n_special = 0, n_special_pad = 0, n_bulk = 0, bulk_end = 0, npad = 0;
  std::vector<int32_t> slot_of;   // reference node -> slot
  std::vector<int32_t> node_of;   // slot -> reference node, or -1
  std::vector<int32_t> nw_slots;  // near-wall list order -> slot
  std::vector<int32_t> inlet_nodes;
  std::vector<uint8_t> kind;      // per reference node

  void* fs[2] = {nullptr, nullptr};
  double* fprev[2] = {nullptr, nullptr};
  int32_t* src = nullptr;
  uint32_t* sflags = nullptr;
  int32_t* bc_nbr = nullptr;
  int32_t* inlet_slot = nullptr;
  double* inlet_basis = nullptr;
  double* normal = nullptr;
  double* rho_u = nullptr;
  double* wss = nullptr;
  float* porosity = nullptr;
  uint32_t* err = nullptr;
  double* f_tmp = nullptr;
  double* partials = nullptr;
  double* red_out = nullptr;
  size_t device_bytes = 0;

  int cur = 0;
  int64_t iter = 0;
  double total_runtime = 0.0;
  bool stepped = false;              // a step ran since the last set/create
  std::vector<double> last_set;      // f given to the last set (node-major), empty = w
  int n_modes = 1;
  int32_t batch_steps = 0;           // steps of the last lbm_step* call (timed by ev0/ev1)
  bool split_timing = false;         // lbm_set_kernel_timing: event pairs around stress flushes
  bool in_batch = false;
  std::vector<hipEvent_t> flush_ev;  // pairs, the first flush_used of them recorded in the last batch
  size_t flush_used = 0;
  int64_t flush_count = 0;           // stress flushes launched in the last batch
  double last_coef[MAX_MODES] = {0};
  StepArgs base{};

  // Column-segment family (half-way bounce-back walls), see columns.h.
  int family = 0;                    // 0: reference rules (kernels.hip), 1: columns.hip
  ChunkMeta* meta = nullptr;
  int32_t* seg_delta = nullptr;
  uint32_t* bbm = nullptr;
  int32_t* irr_src = nullptr;
  uint32_t* irr_kind = nullptr;
  int32_t* irr_bcn = nullptr;
  int32_t* irr_inlet = nullptr;
  int32_t* bslots = nullptr;
  double* nw_normal = nullptr;
  void* nw_normal_c = nullptr;    // (nx, ny, nz, 0) per near-wall node, arithmetic precision
  double* nw_wss = nullptr;
  void* nw_rho_u = nullptr;
  void* nw_ring = nullptr;
  int32_t ring_depth = 8, ring_count = 0;  // stress ring: slots written since the last flush
  int64_t n_irr = 0, n_nw = 0;
  std::vector<int32_t> nw_rows;      // near-wall list order -> near-wall slot rank
  std::vector<int32_t> bsl_slot, bsl_kind, bsl_up, bsl_inlet;  // inlet/outlet list (bslots order)
  std::vector<int32_t> bsl_up_src;  // Q per list entry: raw pull sources of the node behind (-1 = bounce back)
  ColArgs cbase_args{};
  int cache_policy = 0;              // 0 plain, 1 nontemporal stores, 2 alternating hints
  size_t step_cached_bytes = 0;      // metadata and stress data every step reads (kept cached)

  // Reverse-pass workspace, kept across lbm_adjoint_backward calls (freed
  // with the solver): state slots in one arena, the two adjoint fields, and
  // the upstream-node tables of the transposed boundary rules.
  struct AdjWork {
    char* arena = nullptr;
    size_t arena_bytes = 0;
    char* ckpt_arena = nullptr;  // checkpoints of lbm_adjoint_forward
    size_t ckpt_bytes = 0;
    void* rho[2] = {nullptr, nullptr};
    void* mb = nullptr;
    double* own = nullptr;
    double* galpha = nullptr;
    int32_t *tgt = nullptr, *tgt_src = nullptr, *tgt_off = nullptr, *tgt_inl = nullptr, *member_of = nullptr;
    int32_t n_tgt = -1;  // -1: tables not built yet
    char* tape = nullptr;  // moments tape of one checkpoint interval
    size_t tape_bytes = 0;
    uint8_t* spec_flag = nullptr;    // npad: slots kept in the sparse state copies
    int32_t* spec_entries = nullptr;  // slots whose pulls the reverse step re-reads
    size_t spec_cap = 0;
    bool tape_on = true;   // lbm_adjoint_set_tape
  } adjw;

  // Adjoint forward record (lbm_adjoint_forward).
  struct Adjoint {
    bool recorded = false;
    int32_t T = 0, K = 0, M = 0, nck = 0, n_in = 0, n_out = 0;  // M: inlet modes of the record
    std::vector<void*> ckpt;          // h_0, h_K, h_2K, ..., h_T
    void* prev = nullptr;             // h_{T-1} (the input of the last step)
    std::vector<double> coef;         // T * n_modes
    std::vector<double> g0;           // initial f (node-major), empty = w_q
    double wss_weight = 0.0, dp_weight = 0.0;
    double* obj_w = nullptr;          // n_nw
    double* obj_acc = nullptr;        // n_nw
    double* obj_rho = nullptr;        // n_bslots
    uint64_t* obj_mask = nullptr;     // npad / CHUNK: bit l of chunk c marks slot c * CHUNK + l in obj_slots
    std::vector<int32_t> obj_slots;   // slots of the near-wall nodes with a nonzero stress weight
  } adj;
};

namespace {

template <class T>
int dalloc(lbm_solver* s, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return LBM_OK;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  s->device_bytes += count * sizeof(T);
  return LBM_OK;
}

int dalloc_raw(lbm_solver* s, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return LBM_OK;
  HIP_TRY(hipMalloc(p, bytes));
  s->device_bytes += bytes;
  return LBM_OK;
}

void adjoint_release(lbm_solver* s) {
  // The checkpoints live in the solver's persistent arena.
  void* ptrs[] = {s->adj.obj_w, s->adj.obj_acc, s->adj.obj_rho, s->adj.obj_mask};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  s->adj = lbm_solver::Adjoint{};
}

void free_all(lbm_solver* s) {
  adjoint_release(s);
  {
    auto& W = s->adjw;
    void* ptrs[] = {W.arena, W.ckpt_arena, W.rho[0], W.rho[1], W.mb, W.own, W.galpha, W.tgt, W.tgt_src, W.tgt_off,
                    W.tgt_inl, W.member_of, W.tape, W.spec_flag, W.spec_entries};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    W = lbm_solver::AdjWork{};
  }
  void* ptrs[] = {s->fs[0], s->fs[1], s->fprev[0], s->fprev[1], s->src, s->sflags, s->bc_nbr, s->inlet_slot,
                  s->inlet_basis, s->normal, s->rho_u, s->wss, s->porosity, s->err, s->f_tmp, s->partials,
                  s->red_out, s->meta, s->seg_delta, s->bbm, s->irr_src, s->irr_kind, s->irr_bcn,
                  s->irr_inlet, s->bslots, s->nw_normal, s->nw_normal_c, s->nw_wss, s->nw_rho_u, s->nw_ring};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  for (hipEvent_t e : s->flush_ev) (void)hipEventDestroy(e);
  s->flush_ev.clear();
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

StepArgs args_for(lbm_solver* s, int in) {
  StepArgs a = s->base;
  a.fs_in = s->fs[in];
  a.fs_out = s->fs[1 - in];
  a.fprev_in = s->fprev[in];
  a.fprev_out = s->fprev[1 - in];
  return a;
}

constexpr size_t kInfinityCacheBytes = size_t(256) << 20;  // MI355X Infinity Cache (MALL)

ColArgs col_args_for(lbm_solver* s, int in) {
  ColArgs c = s->cbase_args;
  c.fs_in = s->fs[in];
  c.fs_out = s->fs[1 - in];
  // Alternating policy: the step writing fs[1] keeps its output cached and
  // streams its input; the step reading fs[1] streams its output.
  c.cache_hint = s->cache_policy == 0 ? CACHE_PLAIN
                 : s->cache_policy == 1 ? CACHE_NT_STORES
                                        : (in == 0 ? CACHE_NT_LOADS : CACHE_NT_STORES);
  // Alternating hints with room to spare: the first one or two planes of
  // both buffers stay cached as well, so the streamed step reads them from
  // the cache. The budget keeps an 8% margin: the cache is sliced by memory
  // channel, so a fit much tighter than that depends on how the buffers'
  // pages fall on the channels (on the 2.85M-node bifurcation two planes,
  // at 96%, ran 1.7 us faster on most boxes and 9 us slower on one; three,
  // at 99%, thrashed everywhere).
  c.resident_planes = 0;
  if (s->cache_policy == 2) {
    const size_t plane = static_cast<size_t>(s->npad) * s->esize;
    const size_t other = s->step_cached_bytes + (s->porosity ? static_cast<size_t>(s->npad) * 4 : 0);
    for (int k = 2; k >= 1 && !c.resident_planes; --k)
      if (static_cast<size_t>(s->Q + k) * plane + other <= kInfinityCacheBytes / 25 * 23) c.resident_planes = k;
  }
  c.porosity = s->porosity;
  c.n_modes = s->n_modes;
  return c;
}

int ensure_tmp(lbm_solver* s) {
  if (s->f_tmp) return LBM_OK;
  return dalloc(s, &s->f_tmp, static_cast<size_t>(s->Q) * s->npad);
}

// Collide a given f (null = f = w_q everywhere) into the current state.
int init_state(lbm_solver* s, const double* f_host) {
  StepArgs a = args_for(s, 1 - s->cur);  // writes fs[cur], fprev[cur]
  a.f_direct = nullptr;
  if (f_host) {
    int rc = ensure_tmp(s);
    if (rc) return rc;
    std::vector<double> soa(static_cast<size_t>(s->Q) * s->npad, 0.0);
    const lbm::LatticeView L = lbm::lattice_view(s->Q);
    for (int64_t slot = 0; slot < s->npad; ++slot) {
      const int32_t n = s->node_of[slot];
      for (int q = 0; q < s->Q; ++q)
        soa[static_cast<size_t>(q) * s->npad + slot] = n >= 0 ? f_host[int64_t(n) * s->Q + q] : L.w[q];
    }
    HIP_TRY(hipMemcpyAsync(s->f_tmp, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    a.f_direct = s->f_tmp;
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  a.wss_on = 0;
  const bool frc = s->porosity || s->prm.force[0] != 0.0 || s->prm.force[1] != 0.0 || s->prm.force[2] != 0.0;
  if (s->family == 1) {
    ColArgs c = col_args_for(s, 1 - s->cur);
    c.f_direct = a.f_direct;
    HIP_TRY(launch_columns(s->Q, s->prec, KMODE_INIT, frc, c, s->stream));
  } else {
    HIP_TRY(launch_lbm(s->Q, s->prec, KMODE_INIT, frc, a, s->stream));
  }
  s->stepped = false;
  if (f_host) s->last_set.assign(f_host, f_host + static_cast<size_t>(s->N) * s->Q);
  else s->last_set.clear();
  return LBM_OK;
}

bool has_force(const lbm_solver* s) {
  return s->porosity != nullptr || s->prm.force[0] != 0.0 || s->prm.force[1] != 0.0 || s->prm.force[2] != 0.0;
}

// Folds the pending stress ring slots into the FP64 accumulators.
int flush_wss(lbm_solver* s) {
  if (s->family != 1 || s->ring_count == 0) return LBM_OK;
  // The first kTimedFlushes flushes of a batch are timed (their mean stands
  // for all of them; events on every flush would slow the batch itself).
  constexpr size_t kTimedFlushes = 16;
  const bool timed = s->split_timing && s->in_batch && s->flush_used < 2 * kTimedFlushes;
  if (s->in_batch) s->flush_count++;
  if (timed) {
    while (s->flush_ev.size() < s->flush_used + 2) {
      hipEvent_t e = nullptr;
      HIP_TRY(hipEventCreate(&e));
      s->flush_ev.push_back(e);
    }
    HIP_TRY(hipEventRecord(s->flush_ev[s->flush_used], s->stream));
  }
  HIP_TRY(launch_wss_flush(s->prec, s->cbase_args, s->ring_count, s->stream));
  if (timed) {
    HIP_TRY(hipEventRecord(s->flush_ev[s->flush_used + 1], s->stream));
    s->flush_used += 2;
  }
  s->ring_count = 0;
  return LBM_OK;
}

// One column-segment step with the stress ring bookkeeping.
int column_step(lbm_solver* s, ColArgs& c, bool force) {
  if (c.wss_on) c.ring_slot = s->ring_count;
  HIP_TRY(launch_columns(s->Q, s->prec, KMODE_STEP, force, c, s->stream));
  if (c.wss_on && ++s->ring_count == s->ring_depth) return flush_wss(s);
  return LBM_OK;
}

int run_steps(lbm_solver* s, int32_t nsteps, const double* coef) {
  if (nsteps < 0) return fail(LBM_ERR_ARG, "nsteps < 0");
  HIP_TRY(hipSetDevice(s->prm.device));
  // A step leaves the recorded forward pass: its reverse pass would restore
  // the recorded final state under the later iteration count and sums.
  if (nsteps > 0) adjoint_release(s);
  const bool force = has_force(s);
  s->flush_used = 0;
  s->flush_count = 0;
  s->in_batch = true;
  struct BatchEnd {
    lbm_solver* s;
    ~BatchEnd() { s->in_batch = false; }
  } batch_end{s};
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  for (int32_t t = 0; t < nsteps; ++t) {
    StepArgs a = args_for(s, s->cur);
    if (coef) {
      for (int m = 0; m < s->n_modes; ++m) a.coef[m] = coef[int64_t(t) * s->n_modes + m];
    } else {
      const double U = (s->prm.flow_type == LBM_FLOW_PULSATILE) ? lbm_inlet_velocity(s->total_runtime, 0.9)
                                                                 : s->prm.cte_vel;
      for (int m = 0; m < MAX_MODES; ++m) a.coef[m] = 0.0;
      a.coef[0] = U;
    }
    a.wss_on = (s->prm.wss_every > 0 && (s->iter % s->prm.wss_every) == 0) ? 1 : 0;
    if (s->family == 1) {
      ColArgs c = col_args_for(s, s->cur);
      std::memcpy(c.coef, a.coef, sizeof(c.coef));
      c.wss_on = a.wss_on;
      int rc = column_step(s, c, force);
      if (rc) return rc;
    } else {
      HIP_TRY(launch_lbm(s->Q, s->prec, KMODE_STEP, force, a, s->stream));
    }
    std::memcpy(s->last_coef, a.coef, sizeof(s->last_coef));
    s->cur = 1 - s->cur;
    s->iter++;
    if (s->prm.flow_type == LBM_FLOW_PULSATILE) s->total_runtime += s->prm.dt;
    s->stepped = true;
  }
  HIP_TRY(hipEventRecord(s->ev1, s->stream));
  s->batch_steps = nsteps;
  if (s->base.zero_check && nsteps > 0) {
    uint32_t flags = 0;
    HIP_TRY(hipMemcpyAsync(&flags, s->err, sizeof(flags), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (flags) {
      HIP_TRY(hipMemsetAsync(s->err, 0, sizeof(uint32_t), s->stream));
      return fail(LBM_ERR_PARITY,
                  "an f* == 0.0 population was not streamed into a plain fluid node; exact reference "
                  "replay would need its previous populations");
    }
  }
  return LBM_OK;
}

void xcd_tables(int32_t Bs, int32_t Bb, int32_t* first, int32_t* offset) {
  int32_t off = 0;
  for (int x = 0; x < 8; ++x) {
    const int32_t f = Bs + ((x - Bs % 8) + 8) % 8;
    const int32_t cnt = f < Bs + Bb ? (Bs + Bb - 1 - f) / 8 + 1 : 0;
    first[x] = f;
    offset[x] = off;
    off += cnt;
  }
}

// Layout of the column-segment family (columns.h): slots = reference node
// order; regular nodes pull through their segment's 8 lateral deltas,
// irregular nodes through an explicit int32 row.
int create_columns(lbm_solver* s, const Tables& t) {
  const int Q = t.Q;
  const int64_t N = t.N;
  const lbm::LatticeView L = lbm::lattice_view(Q);
  const lbm_params& p = s->prm;
  s->npad = std::max<int64_t>(round_up(N, BLOCK), BLOCK);
  if (s->npad >= (int64_t(1) << 31)) return fail(LBM_ERR_ARG, "too many nodes for int32 slots");
  const int64_t npad = s->npad;
  s->slot_of.resize(N);
  s->node_of.assign(npad, -1);
  lbm::parallel_for(N, 0, [&](int64_t n) {
    s->slot_of[n] = static_cast<int32_t>(n);
    s->node_of[n] = static_cast<int32_t>(n);
  });
  s->n_special = 0;
  s->n_special_pad = 0;
  s->n_bulk = N;
  s->bulk_end = N;
  s->nw_slots = t.nearwall;

  auto srcq = [&](int q, int64_t n) -> int32_t {
    return q == L.rest ? static_cast<int32_t>(n) : t.nb[static_cast<size_t>(L.opp[q]) * N + n];
  };
  const bool inl_b = p.inlet_mode != LBM_INLET_NONE, out_b = p.outlet_mode != LBM_OUTLET_NONE;
  auto bc_node = [&](int64_t n) {
    return (t.kind[n] == KIND_INLET && inl_b) || (t.kind[n] == KIND_OUTLET && out_b);
  };
  // A node whose (i, j) differs from its predecessor's, or whose k does not
  // follow it, starts a column run; every segment lies inside one run.
  auto run_start = [&](int64_t n) {
    if (n == 0) return true;
    const int32_t* c = &t.ijk[3 * static_cast<size_t>(n)];
    return c[0] != c[-3] || c[1] != c[-2] || c[2] != c[-1] + 1;
  };
  // Host ranges of whole runs, one per thread: the encodings below restart
  // at every run, so the ranges are independent.
  const int nth = lbm::resolve_threads(0);
  std::vector<int64_t> bounds{0};
  for (int r = 1; r < nth; ++r) {
    int64_t b = std::max(bounds.back() + 1, N * r / nth);
    while (b < N && !run_start(b)) ++b;
    if (b >= N) break;
    bounds.push_back(b);
  }
  bounds.push_back(N);
  const int nranges = static_cast<int>(bounds.size()) - 1;

  // Column segments: runs of slots with the same (i, j), consecutive k and,
  // for every lateral offset o, one delta src + c_z - n that all existing
  // links of the run's regular nodes share. Walking the slots in order, a
  // node whose links disagree with the deltas its run has established so far
  // (the neighbour column changed segment) starts a new segment; a node that
  // disagrees with itself (a periodic wrap, a link of its own column that is
  // not the next slot) or is an inlet/outlet node takes an explicit row and
  // constrains nothing. Deltas no link establishes stay 0.
  std::vector<int32_t> seg_of(N), delta;
  {
    std::vector<std::vector<int32_t>> part(nranges);
    lbm::parallel_ranges(bounds, [&](int r, int64_t a, int64_t b) {
      std::vector<int32_t>& dl_out = part[r];
      int32_t cur[NCOL];
      bool have[NCOL] = {};
      for (int64_t n = a; n < b; ++n) {
        bool split = n == a || run_start(n);
        bool self_ok = !bc_node(n);
        int32_t nd[NCOL];
        bool nh[NCOL] = {};
        for (int q = 0; q < Q && self_ok; ++q) {
          const int32_t sq = srcq(q, n);
          if (sq < 0) continue;
          const int o = col_of(L.cx[q], L.cy[q]);
          if (o < 0) {
            if (int64_t(sq) != n - L.cz[q]) self_ok = false;
            continue;
          }
          const int32_t dl = static_cast<int32_t>(int64_t(sq) + L.cz[q] - n);
          if (nh[o] && nd[o] != dl) self_ok = false;
          nd[o] = dl;
          nh[o] = true;
        }
        if (self_ok && !split)
          for (int o = 0; o < NCOL; ++o)
            if (nh[o] && have[o] && cur[o] != nd[o]) split = true;
        if (split) {
          dl_out.insert(dl_out.end(), NCOL, 0);
          std::fill(have, have + NCOL, false);
        }
        if (self_ok)
          for (int o = 0; o < NCOL; ++o)
            if (nh[o] && !have[o]) {
              cur[o] = nd[o];
              have[o] = true;
              dl_out[dl_out.size() - NCOL + o] = nd[o];
            }
        seg_of[n] = static_cast<int32_t>(dl_out.size() / NCOL - 1);  // within the range
      }
    });
    std::vector<int64_t> seg_base(nranges + 1, 0);
    for (int r = 0; r < nranges; ++r) seg_base[r + 1] = seg_base[r] + static_cast<int64_t>(part[r].size() / NCOL);
    if (seg_base[nranges] >= (int64_t(1) << 31)) return fail(LBM_ERR_ARG, "too many column segments");
    delta.reserve(static_cast<size_t>(seg_base[nranges]) * NCOL);
    for (auto& v : part) {
      delta.insert(delta.end(), v.begin(), v.end());
      std::vector<int32_t>().swap(v);
    }
    lbm::parallel_ranges(bounds, [&](int r, int64_t a, int64_t b) {
      const int32_t add = static_cast<int32_t>(seg_base[r]);
      for (int64_t n = a; n < b; ++n) seg_of[n] += add;
    });
    if (delta.empty()) delta.assign(NCOL, 0);
  }

  // Regular nodes: every existing link follows its segment's delta (or the
  // own column); missing links become bounce-back bits.
  std::vector<uint8_t> reg(N, 0);
  std::vector<uint32_t> bbm_node(N, 0);
  lbm::parallel_for(N, 0, [&](int64_t n) {
    bool r = !bc_node(n);
    uint32_t bm = 0;
    for (int q = 0; q < Q && r; ++q) {
      const int32_t sq = srcq(q, n);
      if (sq < 0) {
        bm |= 1u << q;
        continue;
      }
      const int o = col_of(L.cx[q], L.cy[q]);
      const int64_t base = o < 0 ? n : n + delta[static_cast<size_t>(seg_of[n]) * NCOL + o];
      if (int64_t(sq) != base - L.cz[q]) r = false;
    }
    reg[n] = r ? 1 : 0;
    bbm_node[n] = r ? bm : 0;
  });

  std::vector<int32_t> inlet_index(N, -1);
  for (size_t i = 0; i < t.inlet_nodes.size(); ++i) inlet_index[t.inlet_nodes[i]] = static_cast<int32_t>(i);
  std::vector<uint8_t> is_nw(N, 0);
  for (int32_t n : t.nearwall) is_nw[n] = 1;

  // Chunk metadata and the compact per-kind arrays, all in slot order: the
  // masks and counts per chunk, their running sums as the chunks' bases,
  // then the entries.
  const int64_t nchunk = npad / CHUNK;
  std::vector<ChunkMeta> meta(nchunk);
  std::vector<int64_t> bsl_base(nchunk + 1, 0);
  lbm::parallel_for(nchunk, 0, [&](int64_t c) {
    ChunkMeta& m = meta[c];
    std::memset(&m, 0, sizeof(m));
    m.seg_first = c * CHUNK < N ? seg_of[c * CHUNK] : 0;
    int32_t nbs = 0;
    for (int l = 0; l < CHUNK; ++l) {
      const int64_t n = c * CHUNK + l;
      if (n >= N) break;
      const uint64_t bit = uint64_t(1) << l;
      if (l > 0 && seg_of[n] != seg_of[n - 1]) m.seg_mask |= bit;
      if (!reg[n]) {
        m.irr_mask |= bit;
        if (bc_node(n)) ++nbs;
      } else if (bbm_node[n]) {
        m.bb_mask |= bit;
      }
      if (is_nw[n]) m.nw_mask |= bit;
    }
    bsl_base[c + 1] = nbs;
  });
  int64_t n_rows = 0, n_bb = 0, n_nwo = 0;
  for (int64_t c = 0; c < nchunk; ++c) {
    ChunkMeta& m = meta[c];
    m.irr_base = static_cast<int32_t>(n_rows);
    m.bb_base = static_cast<int32_t>(n_bb);
    m.nw_base = static_cast<int32_t>(n_nwo);
    n_rows += __builtin_popcountll(m.irr_mask);
    n_bb += __builtin_popcountll(m.bb_mask);
    n_nwo += __builtin_popcountll(m.nw_mask);
    bsl_base[c + 1] += bsl_base[c];
  }
  std::vector<uint32_t> bbm(n_bb);
  std::vector<int32_t> rows(n_rows), nw_slot_order(n_nwo), bsl(bsl_base[nchunk]);
  std::vector<int32_t> nw_rank(N, -1);
  lbm::parallel_for(nchunk, 0, [&](int64_t c) {
    const ChunkMeta& m = meta[c];
    int64_t ir = m.irr_base, ib = m.bb_base, iw = m.nw_base, is = bsl_base[c];
    for (int l = 0; l < CHUNK; ++l) {
      const int64_t n = c * CHUNK + l;
      if (n >= N) break;
      if ((m.irr_mask >> l) & 1u) {
        rows[ir++] = static_cast<int32_t>(n);
        if (bc_node(n)) bsl[is++] = static_cast<int32_t>(n);
      } else if ((m.bb_mask >> l) & 1u) {
        bbm[ib++] = bbm_node[n];
      }
      if ((m.nw_mask >> l) & 1u) {
        nw_rank[n] = static_cast<int32_t>(iw);
        nw_slot_order[iw++] = static_cast<int32_t>(n);
      }
    }
  });
  const int64_t nirr = static_cast<int64_t>(rows.size());
  const int64_t nnw = static_cast<int64_t>(nw_slot_order.size());
  s->n_irr = nirr;
  s->n_nw = nnw;
  s->nw_rows.resize(t.nearwall.size());
  for (size_t i = 0; i < t.nearwall.size(); ++i) s->nw_rows[i] = nw_rank[t.nearwall[i]];
  std::vector<int32_t> isrc(static_cast<size_t>(Q) * std::max<int64_t>(nirr, 1), -1);
  std::vector<uint32_t> ikind(std::max<int64_t>(nirr, 1), 0);
  std::vector<int32_t> ibcn(std::max<int64_t>(nirr, 1), -1), iinl(std::max<int64_t>(nirr, 1), -1);
  lbm::parallel_for(nirr, 0, [&](int64_t r) {
    const int32_t n = rows[r];
    for (int q = 0; q < Q; ++q) isrc[static_cast<size_t>(q) * nirr + r] = srcq(q, n);
    ikind[r] = t.kind[n];
    ibcn[r] = t.bc_nbr[n];
    iinl[r] = inlet_index[n];
  });
  std::vector<double> nrm(3 * static_cast<size_t>(std::max<int64_t>(nnw, 1)), 0.0);
  for (size_t i = 0; i < t.nearwall.size(); ++i)
    for (int d = 0; d < 3; ++d)
      nrm[static_cast<size_t>(d) * nnw + nw_rank[t.nearwall[i]]] = t.normals[3 * i + d];

  // Decode every slot the way the kernel does and compare with the int32
  // pull table (-1 = bounce back).
  std::atomic<int> bad{0};
  lbm::parallel_for(N, 0, [&](int64_t n) {
    const ChunkMeta& m = meta[n / CHUNK];
    const int l = static_cast<int>(n % CHUNK);
    for (int q = 0; q < Q; ++q) {
      int32_t got;
      if ((m.irr_mask >> l) & 1u) {
        got = isrc[static_cast<size_t>(q) * nirr + m.irr_base + rank_below(m.irr_mask, l)];
      } else {
        const uint32_t bm = ((m.bb_mask >> l) & 1u) ? bbm[m.bb_base + rank_below(m.bb_mask, l)] : 0u;
        const int o = col_of(L.cx[q], L.cy[q]);
        const int64_t base = o < 0 ? n : n + delta[static_cast<size_t>(segment_of(m, l)) * NCOL + o];
        got = ((bm >> q) & 1u) ? -1 : static_cast<int32_t>(base - L.cz[q]);
      }
      if (got != srcq(q, n)) bad.store(1, std::memory_order_relaxed);
    }
    if (((m.nw_mask >> l) & 1u) != is_nw[n]) bad.store(2, std::memory_order_relaxed);
  });
  if (bad.load() == 1) return fail(LBM_ERR_STATE, "column-segment encoding does not reproduce the neighbour table");
  if (bad.load() == 2) return fail(LBM_ERR_STATE, "near-wall mask mismatch");

  int rc = LBM_OK;
  const size_t pop_bytes = static_cast<size_t>(Q) * npad * s->esize;
  const size_t ninl = std::max<size_t>(t.inlet_nodes.size(), 1);
  const size_t nnw1 = static_cast<size_t>(std::max<int64_t>(nnw, 1));
  const int64_t nblk = npad / BLOCK;
  if ((rc = dalloc_raw(s, &s->fs[0], pop_bytes)) || (rc = dalloc_raw(s, &s->fs[1], pop_bytes)) ||
      (rc = dalloc(s, &s->meta, meta.size())) || (rc = dalloc(s, &s->seg_delta, delta.size())) ||
      (rc = dalloc(s, &s->bbm, std::max<size_t>(bbm.size(), 1))) ||
      (rc = dalloc(s, &s->irr_src, isrc.size())) || (rc = dalloc(s, &s->irr_kind, ikind.size())) ||
      (rc = dalloc(s, &s->irr_bcn, ibcn.size())) || (rc = dalloc(s, &s->irr_inlet, iinl.size())) ||
      (rc = dalloc(s, &s->bslots, std::max<size_t>(bsl.size(), 1))) ||
      (rc = dalloc(s, &s->nw_normal, 3 * nnw1)) || (rc = dalloc(s, &s->nw_wss, 5 * nnw1)) ||
      (rc = dalloc_raw(s, &s->nw_normal_c, 4 * nnw1 * (s->prec == 0 ? 8 : 4))) ||
      (rc = dalloc_raw(s, &s->nw_rho_u, 4 * nnw1 * (s->prec == 0 ? 8 : 4))) ||
      (rc = dalloc_raw(s, &s->nw_ring, size_t(s->ring_depth) * 4 * nnw1 * (s->prec == 0 ? 8 : 4))) ||
      (rc = dalloc(s, &s->err, 1)) || (rc = dalloc(s, &s->inlet_basis, MAX_MODES * ninl)) ||
      (rc = dalloc(s, &s->partials, static_cast<size_t>(4 * nblk + 2 * ((nnw1 + BLOCK - 1) / BLOCK) + 8))) ||
      (rc = dalloc(s, &s->red_out, 8)))
    return rc;
  std::vector<double> basis(MAX_MODES * ninl, 0.0);
  for (size_t i = 0; i < ninl; ++i) basis[i] = 1.0;
  // AoS copies of the normals in the arithmetic precision.
  std::vector<double> nrm4(4 * nnw1, 0.0);
  for (int64_t i = 0; i < nnw; ++i)
    for (int d = 0; d < 3; ++d) nrm4[4 * i + d] = nrm[static_cast<size_t>(d) * nnw + i];
  const std::vector<float> nrm4_f(nrm4.begin(), nrm4.end());
  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s->stream) : hipSuccess;
  };
  if (up(s->meta, meta.data(), meta.size() * sizeof(ChunkMeta)) || up(s->seg_delta, delta.data(), delta.size() * 4) ||
      up(s->bbm, bbm.data(), bbm.size() * 4) ||
      up(s->irr_src, isrc.data(), isrc.size() * 4) || up(s->irr_kind, ikind.data(), ikind.size() * 4) ||
      up(s->irr_bcn, ibcn.data(), ibcn.size() * 4) || up(s->irr_inlet, iinl.data(), iinl.size() * 4) ||
      up(s->bslots, bsl.data(), bsl.size() * 4) || up(s->nw_normal, nrm.data(), nrm.size() * 8) ||
      (s->prec == 0 ? up(s->nw_normal_c, nrm4.data(), nrm4.size() * 8)
                    : up(s->nw_normal_c, nrm4_f.data(), nrm4_f.size() * 4)) ||
      up(s->inlet_basis, basis.data(), basis.size() * 8) ||
      hipMemsetAsync(s->nw_wss, 0, 5 * nnw1 * 8, s->stream) ||
      hipMemsetAsync(s->nw_rho_u, 0, 4 * nnw1 * (s->prec == 0 ? 8 : 4), s->stream) ||
      hipMemsetAsync(s->err, 0, 4, s->stream) || hipMemsetAsync(s->fs[0], 0, pop_bytes, s->stream) ||
      hipMemsetAsync(s->fs[1], 0, pop_bytes, s->stream))
    return fail(LBM_ERR_HIP, "upload failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(LBM_ERR_HIP, "upload sync failed");

  const double omega = 1.0 / p.tau;
  const double spacing = p.spacing > 0 ? p.spacing : t.grid.spacing;
  const double wss_C = 1060.0 * (p.dt * p.dt) / (spacing * spacing * 1e-6);

  ColArgs& c = s->cbase_args;
  std::memset(&c, 0, sizeof(c));
  c.npad = npad;
  c.n = static_cast<int32_t>(N);
  c.b_blocks = static_cast<int32_t>((bsl.size() + BLOCK - 1) / BLOCK);
  c.m_blocks = static_cast<int32_t>(nblk);
  c.xcd_remap = 1;
  xcd_tables(c.b_blocks, c.m_blocks, c.xcd_first, c.xcd_offset);
  c.meta = s->meta;
  c.seg_delta = s->seg_delta;
  c.bbm = s->bbm;
  c.n_irr = static_cast<int32_t>(nirr);
  c.irr_src = s->irr_src;
  c.irr_kind = s->irr_kind;
  c.irr_bcn = s->irr_bcn;
  c.irr_inlet = s->irr_inlet;
  c.n_nw = static_cast<int32_t>(nnw);
  c.nw_normal = s->nw_normal;
  c.nw_normal_c = s->nw_normal_c;
  c.nw_wss = s->nw_wss;
  c.nw_rho_u = s->nw_rho_u;
  c.nw_ring = s->nw_ring;
  c.wss_C = wss_C;
  c.bslots = s->bslots;
  c.n_bslots = static_cast<int32_t>(bsl.size());
  s->bsl_slot = bsl;
  s->bsl_kind.resize(bsl.size());
  s->bsl_up.resize(bsl.size());
  s->bsl_inlet.resize(bsl.size());
  s->bsl_up_src.assign(bsl.size() * Q, -1);
  for (size_t i = 0; i < bsl.size(); ++i) {
    s->bsl_kind[i] = t.kind[bsl[i]];
    s->bsl_up[i] = t.bc_nbr[bsl[i]];
    s->bsl_inlet[i] = inlet_index[bsl[i]];
    if (s->bsl_up[i] >= 0)
      for (int q = 0; q < Q; ++q) s->bsl_up_src[i * Q + q] = srcq(q, s->bsl_up[i]);
  }
  c.n_inlet = static_cast<int32_t>(ninl);
  c.n_modes = 1;
  c.inlet_basis = s->inlet_basis;
  for (int i = 0; i < 3; ++i) c.inlet_normal[i] = t.inlet_normal[i];
  c.inlet_mode = p.inlet_mode;
  c.outlet_mode = p.outlet_mode;
  c.outlet_rho = p.outlet_rho;
  c.omega = omega;
  for (int i = 0; i < 3; ++i) c.force[i] = p.force[i];
  // Cache policy against the 256 MiB Infinity Cache: both state buffers fit
  // -> plain accesses; one fits -> alternate the hints so that each step's
  // output stays cached for the next step's input (col_args_for); neither ->
  // stream the stores.
  const size_t mall = kInfinityCacheBytes;
  s->cache_policy = 2 * pop_bytes <= mall ? 0 : (pop_bytes <= mall - mall / 8 ? 2 : 1);
  s->step_cached_bytes = meta.size() * sizeof(ChunkMeta) + delta.size() * 4 + bbm.size() * 4 +
                         isrc.size() * 4 + ikind.size() * 12 + nnw1 * 4 * (s->prec == 0 ? 8 : 4) * 2 +
                         basis.size() * 8;
#ifdef LBM_CACHE_POLICY
  s->cache_policy = LBM_CACHE_POLICY;  // tuning builds only
#endif
  // XCD x walks runs x, x + 8, x + 16, ... of 64 blocks (16,384 slots)
  // rather than one contiguous eighth: neighbours along y and z stay in the
  // run (its XCD's L2), and every XCD gets a like share of the geometry, so
  // the XCDs finish together (on the 2.85M-node bifurcation the eighths
  // holding the inclined daughter branches ran 20% longer than the middle
  // ones: FP32 step 86.8 -> 82.1 us, FP64 +6.6%; runs of 16 or 256 blocks
  // gained less or lost). A state whose two buffers both fit the Infinity
  // Cache keeps the contiguous eighths (FP16 on that vessel: 2% faster).
  // FP64 takes runs of 32 blocks (179.0 against 180.2 us on that vessel).
  c.xcd_run = s->cache_policy == 0 ? 0 : (s->prec == 0 ? 32 : 64);
  c.cache_hint = CACHE_PLAIN;

  // Generic fields used by the shared moments kernel and init_state.
  StepArgs& a = s->base;
  std::memset(&a, 0, sizeof(a));
  a.npad = npad;
  a.n_special = 0;
  a.n_special_pad = 0;
  a.bulk_end = static_cast<int32_t>(N);
  a.omega = omega;
  a.err_flags = s->err;
  return LBM_OK;
}

// Column segments of the reference-rule family's slots [0, bulk_end) in slot
// order: runs of slots whose nodes share (i, j) with consecutive k (plain
// fluid in reference order after the special groups, and within each special
// group its nodes in reference order), split where a slot's links disagree
// with the deltas the segment has established (its neighbour column changed,
// or a neighbour sits in another group). One ChunkMeta per 64 slots carries
// seg_first, seg_mask and irr_mask; irr_mask marks every slot for which the
// kernel's decode (lane -> segment from the masks, then p + delta(o) - c_z,
// or p - c_z in its own column) does not give its row of the pull table in
// every direction (a missing link, a padding slot), and those slots read
// their row instead. The decode is replayed here against the table for every
// slot, so a regular slot's sources are its table entries.
void slot_segments(const lbm_solver* s, const Tables& t, const std::vector<int32_t>& src, int Q,
                   std::vector<ChunkMeta>& meta, std::vector<int32_t>& delta) {
  const lbm::LatticeView L = lbm::lattice_view(Q);
  const int64_t npad = s->npad, ns = s->bulk_end;
  const int64_t nchunk = (ns + CHUNK - 1) / CHUNK;
  std::vector<int32_t> seg(static_cast<size_t>(ns));
  delta.clear();
  int32_t cur[NCOL] = {};
  bool have[NCOL] = {};
  for (int64_t p = 0; p < ns; ++p) {
    const int32_t n = s->node_of[p];
    bool split = p == 0 || n < 0 || s->node_of[p - 1] < 0;
    if (!split) {
      const int32_t* c = &t.ijk[3 * static_cast<size_t>(n)];
      const int32_t* b = &t.ijk[3 * static_cast<size_t>(s->node_of[p - 1])];
      split = c[0] != b[0] || c[1] != b[1] || c[2] != b[2] + 1;
    }
    bool ok = n >= 0;
    int32_t nd[NCOL] = {};
    bool nh[NCOL] = {};
    for (int q = 0; q < Q && ok; ++q) {
      const int32_t sq = src[static_cast<size_t>(q) * npad + p];
      if (sq < 0) {
        ok = false;
        break;
      }
      const int o = col_of(L.cx[q], L.cy[q]);
      if (o < 0) {
        if (int64_t(sq) != p - L.cz[q]) ok = false;
        continue;
      }
      const int32_t dl = static_cast<int32_t>(int64_t(sq) + L.cz[q] - p);
      if (nh[o] && nd[o] != dl) ok = false;
      nd[o] = dl;
      nh[o] = true;
    }
    if (ok && !split)
      for (int o = 0; o < NCOL; ++o)
        if (nh[o] && have[o] && cur[o] != nd[o]) split = true;
    if (split) {
      delta.insert(delta.end(), NCOL, 0);
      std::fill(have, have + NCOL, false);
    }
    if (ok)
      for (int o = 0; o < NCOL; ++o)
        if (nh[o] && !have[o]) {
          cur[o] = nd[o];
          have[o] = true;
          delta[delta.size() - NCOL + o] = nd[o];
        }
    seg[static_cast<size_t>(p)] = static_cast<int32_t>(delta.size() / NCOL - 1);
  }
  if (delta.empty()) delta.assign(NCOL, 0);
  meta.assign(static_cast<size_t>(std::max<int64_t>(nchunk, 1)), ChunkMeta{});
  for (int64_t c = 0; c < nchunk; ++c) {
    ChunkMeta& m = meta[c];
    m.seg_first = seg[static_cast<size_t>(c * CHUNK)];
    for (int l = 0; l < CHUNK; ++l) {
      const int64_t p = c * CHUNK + l;
      const uint64_t bit = uint64_t(1) << l;
      if (p >= ns) {
        m.irr_mask |= bit;  // past bulk_end: never stepped
        continue;
      }
      if (l > 0 && seg[p] != seg[p - 1]) m.seg_mask |= bit;
    }
    for (int l = 0; l < CHUNK; ++l) {
      const int64_t p = c * CHUNK + l;
      if (p >= ns) break;
      const int64_t sg = segment_of(m, l);
      bool reg = s->node_of[p] >= 0 && sg == seg[static_cast<size_t>(p)];
      for (int q = 0; q < Q && reg; ++q) {
        const int o = col_of(L.cx[q], L.cy[q]);
        const int64_t dec = (o < 0 ? p : p + delta[static_cast<size_t>(sg) * NCOL + o]) - L.cz[q];
        if (dec != src[static_cast<size_t>(q) * npad + p]) reg = false;
      }
      if (!reg) m.irr_mask |= uint64_t(1) << l;
    }
  }
}

// Device buffers of one reverse pass, released on every exit path.
struct AdjointScratch {
  std::vector<void*> bufs;
  ~AdjointScratch() {
    for (void* p : bufs)
      if (p) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
    if (e == hipSuccess) bufs.push_back(*p);
    return e;
  }
};

}  // namespace

extern "C" {

void lbm_params_default(lbm_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->precision = LBM_PREC_F64;
  p->wall_mode = LBM_WALL_REFERENCE;
  p->inlet_mode = LBM_INLET_REFERENCE_ZOUHE;
  p->outlet_mode = LBM_OUTLET_REFERENCE;
  p->flow_type = LBM_FLOW_CONSTANT;
  p->wss_every = 1;
  p->parity_checks = 1;
  p->tau = 1.0;
  p->dt = 1.0;
  p->outlet_rho = 1.0;
}

int lbm_create(const lbm_geometry* g, const lbm_params* p, lbm_solver** out) {
  if (!g || !p || !out) return fail(LBM_ERR_ARG, "null argument");
  *out = nullptr;
  const Tables& t = g->t;
  if (t.N == 0) return fail(LBM_ERR_ARG, "the geometry has no nodes");
  if (p->precision < 0 || p->precision > 2) return fail(LBM_ERR_ARG, "unknown precision");
  if (!(p->tau > 0.5)) return fail(LBM_ERR_ARG, "tau must exceed 0.5");
  if (p->wall_mode < 0 || p->wall_mode > 1 || p->inlet_mode < 0 || p->inlet_mode > 2 || p->outlet_mode < 0 ||
      p->outlet_mode > 2)
    return fail(LBM_ERR_ARG, "unknown boundary mode");
  if (p->inlet_mode == LBM_INLET_REFERENCE_ZOUHE && !t.inlet_nodes.empty()) {
    if (t.Q != 19)
      return fail(LBM_ERR_UNSUPPORTED,
                  "the reference Zou-He inlet (Boundary.cpp:111-289) is written for D3Q19 only");
    if (!t.has_inlet_normal) return fail(LBM_ERR_ARG, "inlet nodes without an inlet normal");
  }
  if (p->outlet_mode == LBM_OUTLET_REFERENCE)
    for (const auto& lst : t.outlet_nodes)
      for (int32_t n : lst) {
        const int32_t u = t.bc_nbr[n];
        if (u >= 0 && t.kind[u] == KIND_OUTLET)
          return fail(LBM_ERR_UNSUPPORTED,
                      "an outlet node copies from another outlet node; the reference result then depends "
                      "on its sequential processing order");
      }

  auto* s = new lbm_solver();
  s->prm = *p;
  s->Q = t.Q;
  s->prec = p->precision;
  s->esize = p->precision == 0 ? 8 : (p->precision == 1 ? 4 : 2);
  s->N = t.N;
  s->kind = t.kind;
  s->inlet_nodes = t.inlet_nodes;

  auto bail = [&](int rc) {
    free_all(s);
    delete s;
    return rc;
  };
  if (hipSetDevice(p->device) != hipSuccess) return bail(fail(LBM_ERR_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(LBM_ERR_HIP, "stream/event creation failed"));

  if (p->wall_mode == LBM_WALL_HALFWAY_BB) {
    s->family = 1;
    int rc = create_columns(s, t);
    if (rc == LBM_OK) rc = init_state(s, nullptr);
    if (rc == LBM_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(LBM_ERR_HIP, "init failed");
    if (rc) return bail(rc);
    *out = s;
    return LBM_OK;
  }

  // Slot layout: special nodes first, grouped by the rule they run (inlet and
  // outlet nodes, wall nodes, near-wall plain fluid nodes), each group in
  // reference order and padded to whole wavefronts, so that a wavefront runs
  // one rule instead of all of them in turn; then plain fluid in reference
  // order. Slot order is internal: every result is reported per node.
  const int64_t N = t.N;
  std::vector<uint8_t> nearwall(N, 0);
  for (int32_t n : t.nearwall) nearwall[n] = 1;
  s->slot_of.assign(N, -1);
  int64_t ns = 0;
  auto group = [&](auto member) {
    for (int64_t n = 0; n < N; ++n)
      if (member(n)) s->slot_of[n] = static_cast<int32_t>(ns++);
    ns = round_up(ns, 64);
  };
  group([&](int64_t n) { return t.kind[n] == KIND_INLET || t.kind[n] == KIND_OUTLET; });
  group([&](int64_t n) { return t.kind[n] == KIND_WALL; });
  group([&](int64_t n) { return t.kind[n] == KIND_FLUID && nearwall[n]; });
  s->n_special = ns;  // padding slots inside the range carry SFLAG_PAD
  s->n_special_pad = round_up(ns, BLOCK);
  int64_t nb = 0;
  for (int64_t n = 0; n < N; ++n)
    if (s->slot_of[n] < 0) s->slot_of[n] = static_cast<int32_t>(s->n_special_pad + nb++);
  s->n_bulk = nb;
  s->bulk_end = s->n_special_pad + nb;
  s->npad = std::max<int64_t>(round_up(s->bulk_end, BLOCK), BLOCK);
  if (s->npad >= (int64_t(1) << 31)) return bail(fail(LBM_ERR_ARG, "too many nodes for int32 slots"));
  s->node_of.assign(s->npad, -1);
  for (int64_t n = 0; n < N; ++n) s->node_of[s->slot_of[n]] = static_cast<int32_t>(n);
  s->nw_slots.resize(t.nearwall.size());
  for (size_t i = 0; i < t.nearwall.size(); ++i) s->nw_slots[i] = s->slot_of[t.nearwall[i]];

  const int Q = t.Q;
  const lbm::LatticeView L = lbm::lattice_view(Q);
  const int64_t npad = s->npad, sp = s->n_special_pad;

  // Pull table in slot space: source of direction q is the push neighbour in
  // direction opp(q).
  std::vector<int32_t> src(static_cast<size_t>(Q) * npad);
  for (int64_t slot = 0; slot < npad; ++slot) {
    const int32_t n = s->node_of[slot];
    for (int q = 0; q < Q; ++q) {
      int32_t v = static_cast<int32_t>(slot);
      if (n >= 0 && q != L.rest) {
        const int32_t m = t.nb[static_cast<size_t>(L.opp[q]) * N + n];
        v = m >= 0 ? s->slot_of[m] : -1;
      }
      src[static_cast<size_t>(q) * npad + slot] = v;
    }
  }
  std::vector<uint32_t> sflags(sp, SFLAG_PAD);
  std::vector<int32_t> bcn(sp, -1), islot(sp, -1);
  std::vector<double> nrm(3 * sp, 0.0);
  for (int64_t slot = 0; slot < s->n_special; ++slot) {
    const int32_t n = s->node_of[slot];
    if (n < 0) continue;  // padding between groups
    sflags[slot] = uint32_t(t.kind[n]) | (nearwall[n] ? SFLAG_NEARWALL : 0u);
    if (t.bc_nbr[n] >= 0) bcn[slot] = s->slot_of[t.bc_nbr[n]];
  }
  for (size_t i = 0; i < t.inlet_nodes.size(); ++i) islot[s->slot_of[t.inlet_nodes[i]]] = static_cast<int32_t>(i);
  for (size_t i = 0; i < t.nearwall.size(); ++i)
    for (int a = 0; a < 3; ++a) nrm[a * sp + s->slot_of[t.nearwall[i]]] = t.normals[3 * i + a];

  std::vector<ChunkMeta> bmeta;
  std::vector<int32_t> bdelta;
  slot_segments(s, t, src, Q, bmeta, bdelta);

  // Device buffers.
  int rc = LBM_OK;
  const size_t pop_bytes = static_cast<size_t>(Q) * npad * s->esize;
  if ((rc = dalloc_raw(s, &s->fs[0], pop_bytes)) || (rc = dalloc_raw(s, &s->fs[1], pop_bytes)) ||
      (rc = dalloc(s, &s->src, src.size())) || (rc = dalloc(s, &s->meta, bmeta.size())) ||
      (rc = dalloc(s, &s->seg_delta, bdelta.size())) || (rc = dalloc(s, &s->sflags, sp)) || (rc = dalloc(s, &s->bc_nbr, sp)) ||
      (rc = dalloc(s, &s->inlet_slot, sp)) || (rc = dalloc(s, &s->normal, 3 * sp)) ||
      (rc = dalloc(s, &s->rho_u, 4 * sp)) || (rc = dalloc(s, &s->wss, 5 * sp)) || (rc = dalloc(s, &s->err, 1)))
    return bail(rc);
  if (p->wall_mode == LBM_WALL_REFERENCE || p->parity_checks) {
    if ((rc = dalloc(s, &s->fprev[0], static_cast<size_t>(Q) * sp)) ||
        (rc = dalloc(s, &s->fprev[1], static_cast<size_t>(Q) * sp)))
      return bail(rc);
  }
  const int64_t nblk = npad / BLOCK;
  if ((rc = dalloc(s, &s->partials, static_cast<size_t>(4 * nblk + 2 * (sp / BLOCK) + 8))) ||
      (rc = dalloc(s, &s->red_out, 8)))
    return bail(rc);
  const size_t ninl = std::max<size_t>(t.inlet_nodes.size(), 1);
  if ((rc = dalloc(s, &s->inlet_basis, MAX_MODES * ninl))) return bail(rc);
  std::vector<double> basis(MAX_MODES * ninl, 0.0);
  for (size_t i = 0; i < ninl; ++i) basis[i] = 1.0;

  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s->stream) : hipSuccess;
  };
  if (up(s->src, src.data(), src.size() * 4) || up(s->meta, bmeta.data(), bmeta.size() * sizeof(ChunkMeta)) ||
      up(s->seg_delta, bdelta.data(), bdelta.size() * 4) || up(s->sflags, sflags.data(), sp * 4) ||
      up(s->bc_nbr, bcn.data(), sp * 4) || up(s->inlet_slot, islot.data(), sp * 4) ||
      up(s->normal, nrm.data(), nrm.size() * 8) || up(s->inlet_basis, basis.data(), basis.size() * 8) ||
      hipMemsetAsync(s->rho_u, 0, 4 * sp * 8, s->stream) || hipMemsetAsync(s->wss, 0, 5 * sp * 8, s->stream) ||
      hipMemsetAsync(s->err, 0, 4, s->stream) || hipMemsetAsync(s->fs[0], 0, pop_bytes, s->stream) ||
      hipMemsetAsync(s->fs[1], 0, pop_bytes, s->stream))
    return bail(fail(LBM_ERR_HIP, "upload failed"));
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(LBM_ERR_HIP, "upload sync failed"));

  StepArgs& a = s->base;
  std::memset(&a, 0, sizeof(a));
  a.src = s->src;
  a.bmeta = s->meta;
  a.bdelta = s->seg_delta;
  a.npad = npad;
  a.n_special = static_cast<int32_t>(s->n_special);
  a.n_special_pad = static_cast<int32_t>(sp);
  a.bulk_end = static_cast<int32_t>(s->bulk_end);
  a.blocks_special = static_cast<int32_t>(sp / BLOCK);
  a.blocks_bulk = static_cast<int32_t>((s->n_bulk + BLOCK - 1) / BLOCK);
  a.xcd_remap = 1;
  {
    const int32_t Bs = a.blocks_special, Bb = a.blocks_bulk;
    int32_t off = 0;
    for (int x = 0; x < 8; ++x) {
      const int32_t first = Bs + ((x - Bs % 8) + 8) % 8;
      const int32_t cnt = first < Bs + Bb ? (Bs + Bb - 1 - first) / 8 + 1 : 0;
      a.xcd_first[x] = first;
      a.xcd_offset[x] = off;
      off += cnt;
    }
  }
  a.sflags = s->sflags;
  a.bc_nbr = s->bc_nbr;
  a.inlet_slot = s->inlet_slot;
  a.n_inlet = static_cast<int32_t>(ninl);
  a.n_modes = 1;
  a.inlet_basis = s->inlet_basis;
  for (int i = 0; i < 3; ++i) a.inlet_normal[i] = t.inlet_normal[i];
  a.normal = s->normal;
  a.rho_u = s->rho_u;
  a.wss = s->wss;
  // Simulation::ComputeWallShearStress constant (Simulation.cpp:768-771);
  // `spacing` is the doubled lattice pitch (ComplexGeometry.cpp:157).
  {
    const double phys_den = 1060;
    const double dtdt = p->dt * p->dt;
    const double spacing = p->spacing > 0 ? p->spacing : t.grid.spacing;
    const double dxdx = spacing * spacing * 1e-6;
    a.wss_C = phys_den * dtdt / dxdx;
  }
  a.omega = 1.0 / p->tau;
  for (int i = 0; i < 3; ++i) a.force[i] = p->force[i];
  a.porosity = nullptr;
  a.wall_mode = p->wall_mode;
  a.inlet_mode = p->inlet_mode;
  a.outlet_mode = p->outlet_mode;
  a.outlet_rho = p->outlet_rho;
  a.zero_check = (p->parity_checks && p->precision == LBM_PREC_F64 && p->wall_mode == LBM_WALL_REFERENCE) ? 1 : 0;
  // A state that cannot stay in the Infinity Cache is streamed past it, and
  // the pull table stays there (2.85M-node bifurcation, FP64: 367 us per step
  // with plain accesses, 355 with nontemporal population stores, 335 with
  // nontemporal population loads and previous-value stores as well).
  a.nt_stores = pop_bytes > kInfinityCacheBytes ? 1 : 0;
  a.err_flags = s->err;

  if ((rc = init_state(s, nullptr))) return bail(rc);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(LBM_ERR_HIP, "init failed"));
  *out = s;
  return LBM_OK;
}

int lbm_set_inlet_profile(lbm_solver* s, int32_t n_modes, const double* basis) {
  if (!s || n_modes < 1 || n_modes > MAX_MODES || !basis) return fail(LBM_ERR_ARG, "invalid argument");
  const size_t ninl = s->inlet_nodes.size();
  if (ninl == 0) return fail(LBM_ERR_ARG, "no inlet nodes");
  HIP_TRY(hipSetDevice(s->prm.device));
  adjoint_release(s);  // the recorded forward pass used the previous profile
  HIP_TRY(hipMemcpyAsync(s->inlet_basis, basis, n_modes * ninl * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->n_modes = n_modes;
  s->base.n_modes = n_modes;
  return LBM_OK;
}

int lbm_set_porosity(lbm_solver* s, const float* alpha) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  adjoint_release(s);  // the recorded forward pass used the previous field
  // Before the first step the initial collision is redone with the new
  // field, as if it had been given at creation.
  auto recollide = [&]() -> int {
    if (s->stepped) return LBM_OK;
    int rc = init_state(s, s->last_set.empty() ? nullptr : s->last_set.data());
    if (rc == LBM_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(LBM_ERR_HIP, "init failed");
    return rc;
  };
  if (!alpha) {
    if (s->porosity) (void)hipFree(s->porosity);
    s->porosity = nullptr;
    s->base.porosity = nullptr;
    return recollide();
  }
  if (!s->porosity) {
    int rc = dalloc(s, &s->porosity, s->npad);
    if (rc) return rc;
  }
  std::vector<float> h(s->npad, 0.0f);
  for (int64_t slot = 0; slot < s->npad; ++slot)
    if (s->node_of[slot] >= 0) h[slot] = alpha[s->node_of[slot]];
  HIP_TRY(hipMemcpyAsync(s->porosity, h.data(), h.size() * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->base.porosity = s->porosity;
  return recollide();
}

int lbm_step(lbm_solver* s, int32_t nsteps) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  return run_steps(s, nsteps, nullptr);
}

int lbm_step_coefficients(lbm_solver* s, int32_t nsteps, const double* coef) {
  if (!s || !coef) return fail(LBM_ERR_ARG, "null argument");
  return run_steps(s, nsteps, coef);
}

int lbm_step_timed(lbm_solver* s, int32_t nsteps, double* ms) {
  if (!s || !ms) return fail(LBM_ERR_ARG, "null argument");
  int rc = run_steps(s, nsteps, nullptr);
  if (rc) return rc;
  int32_t n = 0;
  return lbm_last_batch_time(s, ms, &n);
}

int lbm_last_batch_time(lbm_solver* s, double* ms, int32_t* nsteps) {
  if (!s || !ms) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  HIP_TRY(hipEventSynchronize(s->ev1));
  float e = 0.0f;
  HIP_TRY(hipEventElapsedTime(&e, s->ev0, s->ev1));
  *ms = e;
  if (nsteps) *nsteps = s->batch_steps;
  return LBM_OK;
}

int lbm_set_wide_offsets(lbm_solver* s, int32_t enable) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  s->cbase_args.wide = enable != 0 ? 1 : 0;
  return LBM_OK;
}

int lbm_set_kernel_timing(lbm_solver* s, int32_t enable) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  s->split_timing = enable != 0;
  return LBM_OK;
}

int lbm_last_batch_kernel_time(lbm_solver* s, double* step_ms, double* other_ms) {
  if (!s || !step_ms || !other_ms) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  HIP_TRY(hipEventSynchronize(s->ev1));
  float e = 0.0f;
  HIP_TRY(hipEventElapsedTime(&e, s->ev0, s->ev1));
  double other = 0.0;
  for (size_t i = 0; i + 1 < s->flush_used; i += 2) {
    float f = 0.0f;
    HIP_TRY(hipEventElapsedTime(&f, s->flush_ev[i], s->flush_ev[i + 1]));
    other += f;
  }
  if (s->flush_used) other *= double(s->flush_count) / double(s->flush_used / 2);
  *other_ms = other;
  *step_ms = double(e) - other;
  return LBM_OK;
}

int lbm_set_populations(lbm_solver* s, const double* f) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  adjoint_release(s);  // a new state ends the recorded forward pass
  int rc = init_state(s, f);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_get_populations(lbm_solver* s, double* f) {
  if (!s || !f) return fail(LBM_ERR_ARG, "null argument");
  const int Q = s->Q;
  if (!s->stepped) {
    if (!s->last_set.empty()) {
      std::memcpy(f, s->last_set.data(), s->last_set.size() * sizeof(double));
    } else {
      const lbm::LatticeView L = lbm::lattice_view(Q);
      for (int64_t n = 0; n < s->N; ++n)
        for (int q = 0; q < Q; ++q) f[n * Q + q] = L.w[q];
    }
    return LBM_OK;
  }
  HIP_TRY(hipSetDevice(s->prm.device));
  int rc = ensure_tmp(s);
  if (rc) return rc;
  // Re-run the last step's stream + boundary stage into f_tmp.
  StepArgs a = args_for(s, 1 - s->cur);
  std::memcpy(a.coef, s->last_coef, sizeof(a.coef));
  a.f_out = s->f_tmp;
  a.wss_on = 0;
  if (s->family == 1) {
    ColArgs c = col_args_for(s, 1 - s->cur);
    std::memcpy(c.coef, s->last_coef, sizeof(c.coef));
    c.f_out = s->f_tmp;
    HIP_TRY(launch_columns(Q, s->prec, KMODE_MATERIALIZE, false, c, s->stream));
  } else {
    HIP_TRY(launch_lbm(Q, s->prec, KMODE_MATERIALIZE, false, a, s->stream));
  }
  std::vector<double> soa(static_cast<size_t>(Q) * s->npad);
  HIP_TRY(hipMemcpyAsync(soa.data(), s->f_tmp, soa.size() * 8, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int64_t n = 0; n < s->N; ++n) {
    const int64_t slot = s->slot_of[n];
    for (int q = 0; q < Q; ++q) f[n * Q + q] = soa[static_cast<size_t>(q) * s->npad + slot];
  }
  return LBM_OK;
}

namespace {
// Post-collision state buffer `buf` (0 or 1) in reference node order, FP64
// unshifted.
int read_state_buffer(lbm_solver* s, int buf, double* fs) {
  HIP_TRY(hipSetDevice(s->prm.device));
  const int Q = s->Q;
  const size_t cnt = static_cast<size_t>(Q) * s->npad;
  std::vector<unsigned char> raw(cnt * s->esize);
  HIP_TRY(hipMemcpyAsync(raw.data(), s->fs[buf], raw.size(), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const lbm::LatticeView L = lbm::lattice_view(Q);
  for (int64_t n = 0; n < s->N; ++n) {
    const size_t slot = static_cast<size_t>(s->slot_of[n]);
    for (int q = 0; q < Q; ++q) {
      const size_t i = static_cast<size_t>(q) * s->npad + slot;
      double v;
      if (s->prec == 0) {
        std::memcpy(&v, raw.data() + 8 * i, 8);
      } else if (s->prec == 1) {
        float x;
        std::memcpy(&x, raw.data() + 4 * i, 4);
        v = double(x) + L.w[q];
      } else {
        uint16_t h;
        std::memcpy(&h, raw.data() + 2 * i, 2);
        v = double(half_to_float(h)) + L.w[q];
      }
      fs[n * Q + q] = v;
    }
  }
  return LBM_OK;
}
}  // namespace

int lbm_get_post_collision(lbm_solver* s, double* fs) {
  if (!s || !fs) return fail(LBM_ERR_ARG, "null argument");
  return read_state_buffer(s, s->cur, fs);
}

int lbm_get_wss(lbm_solver* s, double* wss, double* acc_tau, double* acc_mag) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  if (s->family == 1) {
    int rc = flush_wss(s);
    if (rc) return rc;
    const int64_t m = s->n_nw;
    std::vector<double> h(static_cast<size_t>(5 * m));
    if (m) {
      HIP_TRY(hipMemcpyAsync(h.data(), s->nw_wss, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
    }
    for (size_t i = 0; i < s->nw_rows.size(); ++i) {
      const int64_t r = s->nw_rows[i];
      if (wss) wss[i] = h[r];
      if (acc_tau)
        for (int a = 0; a < 3; ++a) acc_tau[3 * i + a] = h[(1 + a) * m + r];
      if (acc_mag) acc_mag[i] = h[4 * m + r];
    }
    return LBM_OK;
  }
  const int64_t sp = s->n_special_pad;
  std::vector<double> h(static_cast<size_t>(5 * sp));
  if (sp) {
    HIP_TRY(hipMemcpyAsync(h.data(), s->wss, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  for (size_t i = 0; i < s->nw_slots.size(); ++i) {
    const int64_t slot = s->nw_slots[i];
    if (wss) wss[i] = h[slot];
    if (acc_tau)
      for (int a = 0; a < 3; ++a) acc_tau[3 * i + a] = h[(1 + a) * sp + slot];
    if (acc_mag) acc_mag[i] = h[4 * sp + slot];
  }
  return LBM_OK;
}

int lbm_get_osi_tawss(lbm_solver* s, double* osi, double* tawss) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  const size_t n = s->nw_slots.size();
  std::vector<double> acc(3 * n), mag(n);
  int rc = lbm_get_wss(s, nullptr, acc.data(), mag.data());
  if (rc) return rc;
  const double period = 0.9;  // Pulsatile::period (pulsatile.h:53)
  for (size_t i = 0; i < n; ++i) {
    // Simulation::ComputeOSIAndTAWSS (Simulation.cpp:813-824).
    const double t = mag[i] / period;
    const double norm = sqrt(acc[3 * i] * acc[3 * i] + acc[3 * i + 1] * acc[3 * i + 1] + acc[3 * i + 2] * acc[3 * i + 2]);
    if (osi) osi[i] = 0.5 * (1.0 - (norm / mag[i]));
    if (tawss) tawss[i] = t * s->prm.dt;
  }
  return LBM_OK;
}

int lbm_get_macroscopic(lbm_solver* s, double* rho, double* u) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  const int Q = s->Q;
  std::vector<double> f(static_cast<size_t>(s->N) * Q);
  int rc = lbm_get_populations(s, f.data());
  if (rc) return rc;
  const lbm::LatticeView L = lbm::lattice_view(Q);
  for (int64_t n = 0; n < s->N; ++n) {
    const double* fn = f.data() + n * Q;
    double r = 0.0;
    for (int q = 0; q < Q; ++q) r += fn[q];
    double jx = 0.0, jy = 0.0, jz = 0.0;
    for (int q = 0; q < Q; ++q) {
      jx += L.cx[q] * fn[q];
      jy += L.cy[q] * fn[q];
      jz += L.cz[q] * fn[q];
    }
    const bool wall0 = s->kind[n] == KIND_WALL && s->prm.wall_mode == LBM_WALL_REFERENCE;
    if (rho) rho[n] = r;
    if (u) {
      u[3 * n] = wall0 ? 0.0 : jx / r;
      u[3 * n + 1] = wall0 ? 0.0 : jy / r;
      u[3 * n + 2] = wall0 ? 0.0 : jz / r;
    }
  }
  return LBM_OK;
}

}  // extern "C"

namespace {

// lbm_reduce's body over a given workspace: moments -> 4 partial rows per
// block at `part`, stress statistics -> 2 rows per block at `wpart`, the six
// folded values at `res`. Every row and result is written before it is
// read, so the workspace needs no clearing between calls.
int reduce_into(lbm_solver* s, double* part, double* wpart, double* res, lbm_stats* out) {
  StepArgs a = args_for(s, s->cur);
  HIP_TRY(launch_moments(s->Q, s->prec, a, part, res, s->stream));
  if (s->family == 1) {
    int rc = flush_wss(s);
    if (rc) return rc;
  }
  double h[6] = {0, 0, 0, 0, 0, 0};
  const bool any_nw = s->family == 1 ? s->n_nw > 0 : s->n_special_pad > 0;
  if (any_nw) {
    if (s->family == 1) HIP_TRY(launch_columns_wss_stats(s->cbase_args, wpart, res + 4, s->stream));
    else HIP_TRY(launch_wss_stats(a, wpart, res + 4, s->stream));
    HIP_TRY(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(h, res, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::memset(out, 0, sizeof(*out));
  out->mass = h[0];
  out->momentum[0] = h[1];
  out->momentum[1] = h[2];
  out->momentum[2] = h[3];
  out->wss_l2 = sqrt(h[4]);
  out->wss_max = h[5];
  out->iteration = s->iter;
  out->time = s->total_runtime;
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_get_post_collision_moments(lbm_solver* s, const int32_t* nodes, int64_t count, double* rho, double* u) {
  if (!s || count < 0 || (count > 0 && !nodes)) return fail(LBM_ERR_ARG, "invalid argument");
  if (count == 0) return LBM_OK;
  std::vector<int32_t> slots(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; ++i) {
    if (nodes[i] < 0 || nodes[i] >= s->N) return fail(LBM_ERR_ARG, "node index out of range");
    slots[i] = s->slot_of[nodes[i]];
  }
  HIP_TRY(hipSetDevice(s->prm.device));
  int32_t* dslots = nullptr;
  double* dout = nullptr;
  struct Free {
    int32_t** a;
    double** b;
    ~Free() {
      if (*a) (void)hipFree(*a);
      if (*b) (void)hipFree(*b);
    }
  } guard{&dslots, &dout};
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&dslots), slots.size() * sizeof(int32_t)));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&dout), static_cast<size_t>(count) * 4 * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(dslots, slots.data(), slots.size() * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(launch_moments_at(s->Q, s->prec, s->fs[s->cur], s->npad, dslots, count, dout, s->stream));
  std::vector<double> h(static_cast<size_t>(count) * 4);
  HIP_TRY(hipMemcpyAsync(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int64_t i = 0; i < count; ++i) {
    if (rho) rho[i] = h[4 * i];
    if (u)
      for (int a = 0; a < 3; ++a) u[3 * i + a] = h[4 * i + 1 + a];
  }
  return LBM_OK;
}

int lbm_zouhe(int32_t device, int64_t count, double* f, const double* vel, const double normal[3]) {
  if (count < 0 || !normal || (count > 0 && (!f || !vel))) return fail(LBM_ERR_ARG, "invalid argument");
  if (count == 0) return LBM_OK;
  HIP_TRY(hipSetDevice(device));
  double *df = nullptr, *dv = nullptr;
  struct Free {
    double** p[2];
    ~Free() {
      for (double** q : p)
        if (*q) (void)hipFree(*q);
    }
  } guard{{&df, &dv}};
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&df), static_cast<size_t>(count) * 19 * sizeof(double)));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&dv), static_cast<size_t>(count) * sizeof(double)));
  HIP_TRY(hipMemcpy(df, f, static_cast<size_t>(count) * 19 * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dv, vel, static_cast<size_t>(count) * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_zouhe_batch(df, dv, count, normal, nullptr));
  HIP_TRY(hipMemcpy(f, df, static_cast<size_t>(count) * 19 * sizeof(double), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_reduce(lbm_solver* s, lbm_stats* out) {
  if (!s || !out) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  return reduce_into(s, s->partials, s->partials + 4 * (s->npad / BLOCK), s->red_out, out);
}

int lbm_reduce_noalloc(lbm_solver* s, lbm_stats* out) {
  if (!s || !out) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->prm.device));
  double* ws = nullptr;
  int rc = scratch_on(s->prm.device, &ws);
  if (rc) return rc;
  // Same layout as the solver's own workspace: the moments' partial rows
  // (4 per block), the stress statistics' rows (2 per block), the 6 results.
  const int64_t off_wss = 4 * (s->npad / BLOCK);
  const int64_t wss_rows = s->family == 1 ? (s->n_nw + BLOCK - 1) / BLOCK : (s->n_special_pad + BLOCK - 1) / BLOCK;
  const int64_t off_out = round_up(off_wss + 2 * wss_rows, 32);
  const int64_t need = off_out + 8;
  if (need > static_cast<int64_t>(kScratchDoubles))
    return fail(LBM_ERR_ARG, "lbm_reduce_noalloc: this solver's reduction needs " + std::to_string(need * 8) +
                                 " bytes of workspace, more than the " + std::to_string(kScratchDoubles * 8) +
                                 " bytes of its static device array (lbm_reduce has no such limit)");
  return reduce_into(s, ws, ws + off_wss, ws + off_out, out);
}

int lbm_adjoint_forward(lbm_solver* s, int32_t nsteps, const double* coef, const lbm_objective* obj,
                        int32_t checkpoint_every, double* J) {
  if (!s || !obj || !J) return fail(LBM_ERR_ARG, "null argument");
  if (s->family != 1)
    return fail(LBM_ERR_UNSUPPORTED, "the adjoint covers the half-way bounce-back wall rule (LBM_WALL_HALFWAY_BB)");
  if (!(s->prm.inlet_mode == LBM_INLET_NEE_VELOCITY || s->prm.inlet_mode == LBM_INLET_NONE) ||
      !(s->prm.outlet_mode == LBM_OUTLET_NEE_PRESSURE || s->prm.outlet_mode == LBM_OUTLET_NONE))
    return fail(LBM_ERR_UNSUPPORTED, "the adjoint covers the NEE inlet/outlet rules (or none)");
  if (nsteps < 1 || checkpoint_every < 1) return fail(LBM_ERR_ARG, "nsteps and checkpoint_every must be >= 1");
  if (s->stepped)
    return fail(LBM_ERR_STATE, "the adjoint forward pass starts from a state set by lbm_create or lbm_set_populations");
  HIP_TRY(hipSetDevice(s->prm.device));
  adjoint_release(s);
  auto& A = s->adj;
  const int32_t T = nsteps, K = checkpoint_every, M = s->n_modes;
  const size_t state_bytes = static_cast<size_t>(s->Q) * s->npad * s->esize;
  const size_t nnw = static_cast<size_t>(std::max<int64_t>(s->n_nw, 1));
  const size_t nb = std::max<size_t>(s->bsl_slot.size(), 1);

  std::vector<double> w(nnw, 0.0);
  for (size_t i = 0; i < s->nw_rows.size(); ++i)
    w[s->nw_rows[i]] = (obj->wss_mask && !obj->wss_mask[i]) ? 0.0 : obj->wss_weight;
  A.obj_slots.clear();
  for (size_t i = 0; i < s->nw_rows.size(); ++i)
    if (w[s->nw_rows[i]] != 0.0) A.obj_slots.push_back(s->nw_slots[i]);
  // The same set as one bit per slot, read once per wavefront by the reverse
  // step (a per-lane weight load would put a dependent memory round trip in
  // front of every near-wall node's reverse pull).
  std::vector<uint64_t> omask(static_cast<size_t>(s->npad / CHUNK), 0);
  for (int32_t slot : A.obj_slots) omask[slot / CHUNK] |= uint64_t(1) << (slot % CHUNK);
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&A.obj_w), nnw * 8));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&A.obj_acc), nnw * 8));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&A.obj_rho), nb * 8));
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&A.obj_mask), omask.size() * 8));
  HIP_TRY(hipMemcpyAsync(A.obj_mask, omask.data(), omask.size() * 8, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(A.obj_w, w.data(), nnw * 8, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemsetAsync(A.obj_acc, 0, nnw * 8, s->stream));
  HIP_TRY(hipMemsetAsync(A.obj_rho, 0, nb * 8, s->stream));
  for (int32_t k : s->bsl_kind) {
    if (k == KIND_INLET) A.n_in++;
    if (k == KIND_OUTLET) A.n_out++;
  }
  A.wss_weight = obj->wss_weight;
  A.dp_weight = obj->dp_weight;
  // The initial populations move to the record: lbm_get_populations only
  // reads them before the first step, and the forward pass steps.
  A.g0 = std::move(s->last_set);
  s->last_set.clear();
  A.T = T;
  A.K = K;
  A.M = M;
  A.nck = (T + K - 1) / K + 1;
  A.ckpt.assign(A.nck, nullptr);
  // Checkpoints h_0, h_K, ..., h_T and the last step's input h_{T-1}, in a
  // persistent arena (reused by later forward passes of the same size).
  {
    auto& W = s->adjw;
    const size_t need = (static_cast<size_t>(A.nck) + 1) * state_bytes;
    if (W.ckpt_bytes < need) {
      if (W.ckpt_arena) (void)hipFree(W.ckpt_arena);
      W.ckpt_arena = nullptr;
      W.ckpt_bytes = 0;
      if (hipMalloc(reinterpret_cast<void**>(&W.ckpt_arena), need) != hipSuccess) {
        adjoint_release(s);
        return fail(LBM_ERR_HIP, "not enough device memory for the checkpoints");
      }
      W.ckpt_bytes = need;
    }
    for (int32_t i = 0; i < A.nck; ++i) A.ckpt[i] = W.ckpt_arena + static_cast<size_t>(i) * state_bytes;
    A.prev = W.ckpt_arena + static_cast<size_t>(A.nck) * state_bytes;
  }
  A.coef.assign(static_cast<size_t>(T) * M, 0.0);
  HIP_TRY(launch_stream_copy(A.ckpt[0], s->fs[s->cur], state_bytes, s->stream));
  const bool force = has_force(s);
  for (int32_t t = 0; t < T; ++t) {
    ColArgs c = col_args_for(s, s->cur);
    for (int m = 0; m < MAX_MODES; ++m) c.coef[m] = 0.0;
    if (coef) {
      for (int m = 0; m < M; ++m) c.coef[m] = coef[int64_t(t) * M + m];
    } else {
      c.coef[0] = (s->prm.flow_type == LBM_FLOW_PULSATILE) ? lbm_inlet_velocity(s->total_runtime, 0.9) : s->prm.cte_vel;
    }
    for (int m = 0; m < M; ++m) A.coef[int64_t(t) * M + m] = c.coef[m];
    c.wss_on = 1;
    c.obj_acc = A.obj_acc;
    c.obj_w = A.obj_w;
    c.obj_rho = A.obj_rho;
    int rc = column_step(s, c, force);
    if (rc) return rc;
    std::memcpy(s->last_coef, c.coef, sizeof(s->last_coef));
    s->cur = 1 - s->cur;
    s->iter++;
    if (s->prm.flow_type == LBM_FLOW_PULSATILE) s->total_runtime += s->prm.dt;
    s->stepped = true;
    if ((t + 1) % K == 0 || t + 1 == T)
      HIP_TRY(launch_stream_copy(A.ckpt[(t + 1 + K - 1) / K], s->fs[s->cur], state_bytes, s->stream));
  }
  HIP_TRY(launch_stream_copy(A.prev, s->fs[1 - s->cur], state_bytes, s->stream));
  std::vector<double> acc(nnw), rho(nb);
  HIP_TRY(hipMemcpyAsync(acc.data(), A.obj_acc, nnw * 8, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(rho.data(), A.obj_rho, nb * 8, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  double sw = 0.0, rin = 0.0, rout = 0.0;
  for (size_t i = 0; i < static_cast<size_t>(s->n_nw); ++i) sw += acc[i];
  for (size_t i = 0; i < s->bsl_slot.size(); ++i) {
    if (s->bsl_kind[i] == KIND_INLET) rin += rho[i];
    if (s->bsl_kind[i] == KIND_OUTLET) rout += rho[i];
  }
  double dp = 0.0;
  if (A.n_in) dp += rin / A.n_in;
  if (A.n_out) dp -= rout / A.n_out;
  *J = (sw + A.dp_weight * dp / 3.0) / T;
  A.recorded = true;
  return LBM_OK;
}

int lbm_adjoint_backward(lbm_solver* s, int32_t slots, double* grad_coef, double* grad_porosity) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  auto& A = s->adj;
  if (!A.recorded) return fail(LBM_ERR_STATE, "no adjoint forward pass recorded");
  HIP_TRY(hipSetDevice(s->prm.device));
  const int Q = s->Q;
  const int32_t T = A.T, K = A.K, M = A.M;
  const int64_t npad = s->npad;
  const size_t state_bytes = static_cast<size_t>(Q) * npad * s->esize;
  const size_t aes = s->prec == 0 ? 8 : 4;  // adjoint arithmetic: FP64 or FP32
  const size_t nnw = static_cast<size_t>(std::max<int64_t>(s->n_nw, 1));
  const size_t nb = std::max<size_t>(s->bsl_slot.size(), 1);
  ColArgs cb = col_args_for(s, 0);
  const int32_t n_inlet = cb.n_inlet;

  AdjointScratch S;  // per-call buffers (their sizes depend on the forward pass)
  auto& W = s->adjw;
  auto keep = [&](auto** p, size_t bytes) -> hipError_t {
    if (*p) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(bytes, 8));
  };
  HIP_TRY(keep(&W.rho[0], static_cast<size_t>(Q) * npad * aes));
  HIP_TRY(keep(&W.rho[1], static_cast<size_t>(Q) * npad * aes));
  HIP_TRY(keep(&W.own, nnw * Q * 8));
  HIP_TRY(keep(&W.mb, nb * (Q + 1) * aes));
  HIP_TRY(keep(&W.galpha, static_cast<size_t>(npad) * 8));
  double *gamma = nullptr, *g0 = nullptr, *gcoef = nullptr;
  HIP_TRY(S.alloc(&gamma, static_cast<size_t>(T) * n_inlet * 8));
  HIP_TRY(S.alloc(&gcoef, static_cast<size_t>(T) * std::max(M, 1) * 8));
  HIP_TRY(hipMemsetAsync(W.galpha, 0, static_cast<size_t>(npad) * 8, s->stream));
  HIP_TRY(hipMemsetAsync(gamma, 0, static_cast<size_t>(T) * n_inlet * 8, s->stream));
  if (!A.g0.empty()) {
    std::vector<double> soa(static_cast<size_t>(Q) * npad, 0.0);
    const lbm::LatticeView L = lbm::lattice_view(Q);
    for (int64_t slot = 0; slot < npad; ++slot) {
      const int32_t n = s->node_of[slot];
      for (int q = 0; q < Q; ++q) soa[static_cast<size_t>(q) * npad + slot] = n >= 0 ? A.g0[int64_t(n) * Q + q] : L.w[q];
    }
    HIP_TRY(S.alloc(&g0, soa.size() * 8));
    HIP_TRY(hipMemcpyAsync(g0, soa.data(), soa.size() * 8, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }

  // Upstream nodes of extrapolated inlet/outlet nodes, grouped (CSR), with
  // their raw pull sources and inlet indices (built once per solver).
  if (W.n_tgt < 0) {
    std::vector<std::pair<int32_t, int32_t>> up;  // (upstream slot, bslots index)
    for (size_t i = 0; i < s->bsl_slot.size(); ++i) {
      const int32_t u = s->bsl_up[i], k = s->bsl_kind[i];
      const bool nee = u >= 0 && ((k == KIND_INLET && s->prm.inlet_mode == LBM_INLET_NEE_VELOCITY) ||
                                  (k == KIND_OUTLET && s->prm.outlet_mode == LBM_OUTLET_NEE_PRESSURE));
      if (nee) up.emplace_back(s->slot_of[u], static_cast<int32_t>(i));
    }
    std::sort(up.begin(), up.end());
    std::vector<int32_t> h_tgt, h_off, h_b, h_inl, h_src;
    for (size_t i = 0; i < up.size(); ++i) {
      if (i == 0 || up[i].first != up[i - 1].first) {
        h_tgt.push_back(up[i].first);
        h_off.push_back(static_cast<int32_t>(i));
      }
      h_b.push_back(up[i].second);
      h_inl.push_back(s->bsl_kind[up[i].second] == KIND_INLET ? s->bsl_inlet[up[i].second] : -1);
    }
    h_off.push_back(static_cast<int32_t>(up.size()));
    h_src.resize(h_tgt.size() * Q);
    for (size_t k = 0; k < h_tgt.size(); ++k)
      for (int q = 0; q < Q; ++q) h_src[k * Q + q] = s->bsl_up_src[size_t(h_b[h_off[k]]) * Q + q];
    std::vector<int32_t> h_member(std::max<size_t>(s->bsl_slot.size(), 1), -1);
    for (size_t j = 0; j < h_b.size(); ++j) h_member[h_b[j]] = static_cast<int32_t>(j);
    HIP_TRY(keep(&W.tgt, h_tgt.size() * 4));
    HIP_TRY(keep(&W.tgt_off, h_off.size() * 4));
    HIP_TRY(keep(&W.tgt_inl, h_inl.size() * 4));
    HIP_TRY(keep(&W.tgt_src, h_src.size() * 4));
    HIP_TRY(keep(&W.member_of, h_member.size() * 4));
    HIP_TRY(hipMemcpyAsync(W.member_of, h_member.data(), h_member.size() * 4, hipMemcpyHostToDevice, s->stream));
    if (!h_tgt.empty()) {
      HIP_TRY(hipMemcpyAsync(W.tgt, h_tgt.data(), h_tgt.size() * 4, hipMemcpyHostToDevice, s->stream));
      HIP_TRY(hipMemcpyAsync(W.tgt_inl, h_inl.data(), h_inl.size() * 4, hipMemcpyHostToDevice, s->stream));
      HIP_TRY(hipMemcpyAsync(W.tgt_src, h_src.data(), h_src.size() * 4, hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipMemcpyAsync(W.tgt_off, h_off.data(), h_off.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    W.n_tgt = static_cast<int32_t>(h_tgt.size());
  }

  // Moments tape: while recomputing a checkpoint interval, the step kernel
  // also stores the moments of every slot's pulled populations (4 values
  // instead of 19), so the reverse step of each recomputed step reads them
  // instead of re-pulling its input state. Steps that the recomputation
  // does not execute (the last one of each interval) re-pull. Taken only
  // where it fits in a quarter of the free memory.
  const size_t tape_step = static_cast<size_t>(4) * npad * aes;
  char* tape = nullptr;
  if (W.tape_on && s->family == 1) {
    const size_t tneed = static_cast<size_t>(K) * tape_step;
    if (W.tape_bytes < tneed) {
      if (W.tape) (void)hipFree(W.tape);
      W.tape = nullptr;
      W.tape_bytes = 0;
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      if (tneed <= free_b / 4) {
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&W.tape), tneed));
        W.tape_bytes = tneed;
      }
    }
    tape = W.tape;
  }
  std::vector<char> tape_ok(static_cast<size_t>(std::max(K, 1)), 0);
  int32_t l_cur = 0;  // first step of the interval being reversed

  // State buffers for the binomial schedule: `slots` of them plus a
  // scratch state, in one arena that later calls reuse.
  const int32_t max_slots = std::max<int32_t>(K - 1, 0);
  if (slots <= 0 || slots > max_slots) {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t reserve = size_t(1) << 30;
    const size_t avail = free_b + W.arena_bytes;
    const int64_t fit = avail > reserve ? static_cast<int64_t>((avail - reserve) / state_bytes) - 1 : 0;
    const int32_t want = slots <= 0 ? max_slots : std::min(slots, max_slots);
    slots = static_cast<int32_t>(std::max<int64_t>(0, std::min<int64_t>(want, fit)));
  }
  const size_t need = (static_cast<size_t>(slots) + 1) * state_bytes;
  if (W.arena_bytes < need) {
    if (W.arena) (void)hipFree(W.arena);
    W.arena = nullptr;
    W.arena_bytes = 0;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&W.arena), need));
    W.arena_bytes = need;
  }
  std::vector<void*> slot_buf(slots, nullptr);
  for (int32_t i = 0; i < slots; ++i) slot_buf[i] = W.arena + static_cast<size_t>(i) * state_bytes;
  void* scratch = W.arena + static_cast<size_t>(slots) * state_bytes;

  const bool force = has_force(s);
  ColArgs cf = cb;  // forward recomputation: no stress, no objective
  cf.cache_hint = s->cache_policy == 0 ? CACHE_PLAIN : CACHE_NT_STORES;
  cf.resident_planes = 0;
  cf.wss_on = 0;
  cf.obj_acc = nullptr;
  cf.obj_w = nullptr;
  cf.obj_rho = nullptr;
  auto advance = [&](void* src, int32_t from, int32_t to, void* dst) -> int {
    for (int32_t k = from; k < to; ++k) {
      void* out = (k == to - 1) ? dst : s->fs[(k - from) & 1];
      ColArgs c = cf;
      c.fs_in = src;
      c.fs_out = out;
      for (int m = 0; m < MAX_MODES; ++m) c.coef[m] = m < M ? A.coef[int64_t(k) * M + m] : 0.0;
      if (tape) {  // step k + 1 of the interval: tape entry k - l_cur
        c.tape = tape + static_cast<size_t>(k - l_cur) * tape_step;
        HIP_TRY(launch_columns(Q, s->prec, KMODE_TAPE, force, c, s->stream));
        tape_ok[k - l_cur] = 1;
      } else {
        HIP_TRY(launch_columns(Q, s->prec, KMODE_STEP, force, c, s->stream));
      }
      src = out;
    }
    return LBM_OK;
  };
  AdjArgs a{};
  a.c = cb;
  a.c.obj_acc = nullptr;
  a.c.obj_rho = nullptr;
  a.own = W.own;
  a.mb = W.mb;
  a.g_alpha = W.galpha;
  a.inv_T = 1.0 / T;
  a.dp_in = A.n_in ? A.dp_weight / (3.0 * T * A.n_in) : 0.0;
  a.dp_out = A.n_out ? A.dp_weight / (3.0 * T * A.n_out) : 0.0;
  a.obj_w = A.obj_w;
  a.obj_mask = A.obj_mask;
  a.n_tgt = W.n_tgt;
  a.tgt = W.tgt;
  a.tgt_src = W.tgt_src;
  a.tgt_off = W.tgt_off;
  a.tgt_inl = W.tgt_inl;
  a.member_of = W.member_of;
  void* rho[2] = {W.rho[0], W.rho[1]};
  auto reverse = [&](int32_t t, void* h) -> int {
    AdjArgs x = a;
    x.mode = ADJ_STEP;
    x.c.fs_in = h;
    for (int m = 0; m < MAX_MODES; ++m) x.c.coef[m] = m < M ? A.coef[int64_t(t - 1) * M + m] : 0.0;
    x.rho_in = t == T ? nullptr : rho[(t + 1) & 1];
    x.rho_out = rho[t & 1];
    x.tape_in = (tape && tape_ok[t - 1 - l_cur]) ? tape + static_cast<size_t>(t - 1 - l_cur) * tape_step : nullptr;
    // Alternate the cache hints so that every other launch reads an adjoint
    // field still held in the Infinity Cache: odd steps write rho[1] with
    // cached stores (and stream their input), even steps read it plainly
    // (and stream their output).
    x.rho_cached = (t & 1) ? 0 : 1;
    x.gamma = gamma + int64_t(t - 1) * n_inlet;
    HIP_TRY(launch_adjoint(Q, s->prec, x, s->stream));
    HIP_TRY(launch_adjoint_boundary(Q, s->prec, x, s->stream));
    return LBM_OK;
  };
  // Sparse path: with the tape and a buffer for every step of an interval,
  // the recomputation runs through each interval once on the two state
  // buffers, with the forward step's alternating cache hints, and keeps of
  // each state only the slots the reverse step still reads from it (pull
  // sources and own slots of objective and inlet/outlet nodes).
  const bool sparse = tape && max_slots > 0 && slots >= max_slots;
  if (sparse) {
    std::vector<int32_t> ent(A.obj_slots);
    for (size_t i = 0; i < s->bsl_slot.size(); ++i) {
      const int32_t u = s->bsl_up[i], k = s->bsl_kind[i];
      const bool nee = u >= 0 && ((k == KIND_INLET && s->prm.inlet_mode == LBM_INLET_NEE_VELOCITY) ||
                                  (k == KIND_OUTLET && s->prm.outlet_mode == LBM_OUTLET_NEE_PRESSURE));
      ent.push_back(s->bsl_slot[i]);
      if (nee) ent.push_back(s->slot_of[u]);
    }
    if (W.spec_cap < ent.size()) {
      if (W.spec_entries) (void)hipFree(W.spec_entries);
      W.spec_entries = nullptr;
      W.spec_cap = 0;
      HIP_TRY(hipMalloc(reinterpret_cast<void**>(&W.spec_entries), ent.size() * 4));
      W.spec_cap = ent.size();
    }
    HIP_TRY(keep(&W.spec_flag, static_cast<size_t>(npad)));
    if (!ent.empty())
      HIP_TRY(hipMemcpyAsync(W.spec_entries, ent.data(), ent.size() * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(W.spec_flag, 0, static_cast<size_t>(npad), s->stream));
    HIP_TRY(launch_spec_flags(Q, s->prec, cb, W.spec_entries, static_cast<int32_t>(ent.size()), W.spec_flag,
                              s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));  // ent is freed on return
  }
  for (int32_t j = A.nck - 2; j >= 0; --j) {
    const int32_t l = j * K, n = std::min(K, T - l);
    l_cur = l;
    std::fill(tape_ok.begin(), tape_ok.end(), 0);
    if (sparse) {
      void* src = A.ckpt[j];
      for (int32_t k = l; k + 1 < l + n; ++k) {
        void* out = s->fs[(k - l) & 1];
        ColArgs c = cf;
        c.fs_in = src;
        c.fs_out = out;
        for (int m = 0; m < MAX_MODES; ++m) c.coef[m] = m < M ? A.coef[int64_t(k) * M + m] : 0.0;
        if (s->cache_policy == 2) c.cache_hint = out == s->fs[1] ? CACHE_NT_LOADS : CACHE_NT_STORES;
        c.tape = tape + static_cast<size_t>(k - l) * tape_step;
        c.spec_out = slot_buf[k - l];
        c.spec_flag = W.spec_flag;
        HIP_TRY(launch_columns(Q, s->prec, KMODE_TAPE, force, c, s->stream));
        tape_ok[k - l] = 1;
        src = out;
      }
      // h_{t-1}: the checkpoint, the last recomputed state in full (its step
      // has no tape: the recomputation does not execute it), or a sparse copy.
      for (int32_t t = l + n; t > l; --t) {
        void* h = (t - 1 == l) ? A.ckpt[j] : (t == l + n ? src : slot_buf[t - 2 - l]);
        const int rc = reverse(t, h);
        if (rc) return rc;
      }
      continue;
    }
    auto buffer = [&](int32_t id) -> void* {
      return id == REV_BASE ? A.ckpt[j] : (id == REV_SCRATCH ? scratch : slot_buf[id]);
    };
    for (const RevAction& act : revolve_schedule(l, n, slots)) {
      int rc = act.op == REV_ADVANCE ? advance(buffer(act.buf_from), act.from, act.to, buffer(act.buf_to))
                                     : reverse(act.to, buffer(act.buf_from));
      if (rc) return rc;
    }
  }
  {
    AdjArgs x = a;  // lambda_0 and the initial collision's porosity term
    x.mode = ADJ_FINAL;
    x.c.fs_in = A.ckpt[0];
    x.rho_in = rho[1];
    x.rho_out = nullptr;
    x.g0 = g0;
    HIP_TRY(launch_adjoint(Q, s->prec, x, s->stream));
  }
  if (grad_coef) {
    HIP_TRY(launch_coef_gradient(gamma, cb.inlet_basis, T, n_inlet, M, gcoef, s->stream));
    HIP_TRY(hipMemcpyAsync(grad_coef, gcoef, static_cast<size_t>(T) * M * 8, hipMemcpyDeviceToHost, s->stream));
  }
  std::vector<double> ga;
  if (grad_porosity) {
    ga.resize(npad);
    HIP_TRY(hipMemcpyAsync(ga.data(), W.galpha, static_cast<size_t>(npad) * 8, hipMemcpyDeviceToHost, s->stream));
  }
  // Restore the forward pass's last two states (the current state and the
  // input of the last step, which lbm_get_populations re-streams).
  HIP_TRY(hipMemcpyAsync(s->fs[s->cur], A.ckpt[A.nck - 1], state_bytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->fs[1 - s->cur], A.prev, state_bytes, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (grad_porosity)
    for (int64_t n = 0; n < s->N; ++n) grad_porosity[n] = ga[s->slot_of[n]];
  return LBM_OK;
}

int lbm_adjoint_set_tape(lbm_solver* s, int32_t enable) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  s->adjw.tape_on = enable != 0;
  return LBM_OK;
}

int lbm_write_output(const lbm_geometry* g, lbm_solver* s, int32_t iter, const char* phase_name, const char* dir,
                     int32_t pulsatile) {
  if (!g || !s || !phase_name || !dir) return fail(LBM_ERR_ARG, "null argument");
  const lbm::Tables& t = g->t;
  if (t.N != s->N || t.Q != s->Q) return fail(LBM_ERR_ARG, "geometry does not match the solver");
  // WriteOutput(iter) follows iteration iter, whose Collision set each
  // node's density and velocity from f(iter) (Simulation.cpp:77-139; wall
  // nodes carry zero velocity in the reference wall rule). A step here
  // streams f*(iter) into f(iter + 1) and then already collides that, so
  // f*(iter) is the previous post-collision buffer, whose moments BGK keeps
  // equal to those of f(iter) (to rounding). Before any step since the state
  // was set, that buffer holds no collision.
  if (!s->stepped) return fail(LBM_ERR_ARG, "no step since the state was set: output follows an iteration");
  const int Q = s->Q;
  const int64_t N = s->N;
  std::vector<double> f(static_cast<size_t>(N) * Q);
  int rc = read_state_buffer(s, 1 - s->cur, f.data());
  if (rc) return rc;
  const lbm::LatticeView L = lbm::lattice_view(Q);
  std::vector<double> rho(N), u(3 * N), vm(N);
  for (int64_t n = 0; n < N; ++n) {
    const double* fn = f.data() + n * Q;
    double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    for (int q = 0; q < Q; ++q) {
      r += fn[q];
      jx += L.cx[q] * fn[q];
      jy += L.cy[q] * fn[q];
      jz += L.cz[q] * fn[q];
    }
    rho[n] = r;
    const bool zero = s->prm.wall_mode == LBM_WALL_REFERENCE && t.kind[n] == LBM_KIND_WALL;
    u[3 * n] = zero ? 0.0 : jx / r;
    u[3 * n + 1] = zero ? 0.0 : jy / r;
    u[3 * n + 2] = zero ? 0.0 : jz / r;
    vm[n] = std::sqrt(u[3 * n] * u[3 * n] + u[3 * n + 1] * u[3 * n + 1] + u[3 * n + 2] * u[3 * n + 2]);
  }
  const size_t m = t.nearwall.size();
  std::vector<double> wss(m), osi(pulsatile ? m : 0), tawss(pulsatile ? m : 0), nvm(m);
  if ((rc = lbm_get_wss(s, wss.data(), nullptr, nullptr))) return rc;
  if (pulsatile && (rc = lbm_get_osi_tawss(s, osi.data(), tawss.data()))) return rc;
  // The CSV's VelMag is the node's velMag, which only the D3Q19 branch of
  // Collision sets (Simulation.cpp:136-139); it stays at its initial 0 on
  // D3Q27 (Node.h:64).
  for (size_t i = 0; i < m; ++i) nvm[i] = Q == 19 ? vm[t.nearwall[i]] : 0.0;
  // Simulation::WriteOutput (Simulation.cpp:835-856), relative to `dir`.
  const std::string base(dir), it = std::to_string(iter);
  for (const char* sub : {"/Fluid_Variables", "/FlowHemodynamics"})
    if (mkdir((base + sub).c_str(), 0777) != 0 && errno != EEXIST) return fail(LBM_ERR_IO, "cannot create " + base + sub);
  std::string err;
  const std::string raw = base + "/" + it + "_FlowHemodynamics.csv";
  if (!lbm::write_vtu(t, rho, u, base + "/Fluid_Variables/" + it + ".vtu", &err) ||
      !lbm::write_surface_csv(t, wss, osi, tawss, nvm, pulsatile != 0, raw, &err) ||
      !lbm::smooth_surface_csv(raw, base + "/FlowHemodynamics/" + std::string(phase_name) + "_" + it +
                                        "_FlowHemodynamics.csv",
                               0.25, pulsatile != 0, &err))
    return fail(LBM_ERR_IO, err);
  return LBM_OK;
}

int lbm_synchronize(lbm_solver* s) {
  if (!s) return fail(LBM_ERR_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_get_solver_info(const lbm_solver* s, lbm_solver_info* o) {
  if (!s || !o) return fail(LBM_ERR_ARG, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->n_nodes = s->N;
  o->n_special = s->N - s->n_bulk;
  o->n_bulk = s->n_bulk;
  if (s->prm.wall_mode == LBM_WALL_HALFWAY_BB) {  // column segments: explicit rows vs segment-decoded slots
    o->n_special = s->n_irr;
    o->n_bulk = s->N - s->n_irr;
  }
  o->n_pad = s->npad;
  o->device_bytes = static_cast<int64_t>(s->device_bytes);
  o->q = s->Q;
  o->precision = s->prec;
  o->bytes_per_update = 2.0 * s->Q * double(s->esize);
  return LBM_OK;
}

void lbm_destroy(lbm_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->prm.device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
}

}  // extern "C"
