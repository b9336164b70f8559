// Discrete adjoint of the column-segment step (columns.h): argument block
// and launchers.
//
// Forward step t (input h = f*_{t-1}, the stored post-collision state):
//   R_t(x)  raw pull of h at x (half-way bounce-back on missing links)
//   G_t(x)  = R_t(x), or at an inlet/outlet node b with an upstream node u_b
//             the non-equilibrium extrapolation NEE_b(R_t(u_b), U_b(t))
//   f*_t(x) = C(G_t(x); alpha(x))          BGK + Guo force + Brinkman term
//   J       = (1/T) sum_t [ sum_{wall segment} w |tau_t|^2 + a_p dp_t ]
// Reverse step t, with lambda_t = dJ/df*_t:
//   mu_t(x)  = C'(G_t(x))^T lambda_t(x) + dJ_t/dG_t(x)
//   rho_t(y) = mu_t(y) (non-extrapolated nodes) + sum_{b: u_b = y} NEE_b'^T mu_t(b)
//   lambda_{t-1}(z)[q] = rho_t[q](z + c_q), or rho_t[opp q](z) where that link
//                        is missing (transposed bounce-back)   ("reverse pull")
//                        + dJ_t/dh(z) (the stress equilibrium of near-wall z)
// Kernel A(t) forms lambda_t by the reverse pull of rho_{t+1} (plus the
// stress-equilibrium term of step t+1), recomputes G_t from h_{t-1}, and
// writes rho_t without the extrapolation terms; kernel B(t) adds those and
// the inlet-velocity gradient of step t.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "columns.h"

namespace lbm {

constexpr int ADJ_STEP = 0;   // reverse step t (t = T .. 1)
constexpr int ADJ_FINAL = 1;  // lambda_0 and the initial collision's porosity term

struct AdjArgs {
  ColArgs c;             // layout and forward parameters; c.fs_in = h_{t-1}; c.coef = step t's
  int32_t mode;
  const void* rho_in;    // rho_{t+1} (adjoint precision, Q * npad), null at t = T
  void* rho_out;         // rho_t
  int32_t rho_cached;    // 1: rho_in was written with cached stores (read it plainly, stream rho_out);
                         // 0: read rho_in streamed, keep rho_out cached
  const void* tape_in;   // moments (s, jx, jy, jz) of step t's pulled populations, written by the
                         // recomputation (ColArgs::tape layout), or null: re-pull them from h_{t-1}
  double* own;           // n_nw * Q: stress-equilibrium term (read: step t+1's, written: step t's)
  void* mb;              // (Q + 1) per member (CSR order below): mu_t of an extrapolated
                         // inlet/outlet node and its inlet speed U_b(t)
  double* g_alpha;       // npad: dJ/dalpha accumulated over steps
  double inv_T;
  double dp_in, dp_out;  // a_p / (3 T n_in), a_p / (3 T n_out)
  const double* obj_w;   // n_nw: stress weight (0 = outside the wall segment)
  const uint64_t* obj_mask;  // npad / CHUNK: bit l of chunk c set = slot c * CHUNK + l has a nonzero weight
  const double* g0;      // ADJ_FINAL: initial populations (FP64 unshifted, q * npad + slot), null = w_q
  // Kernel B: upstream nodes of extrapolated boundary nodes (CSR), with
  // everything the transposed rules need resolved on the host.
  int32_t n_tgt;
  const int32_t* tgt;      // target slot
  const int32_t* tgt_src;  // Q * n_tgt: raw pull source of the target (-1 = bounce back)
  const int32_t* tgt_off;  // n_tgt + 1
  const int32_t* tgt_inl;  // members: inlet basis index, or -1 for an outlet node
  const int32_t* member_of;  // n_bslots: member position of an extrapolated node, or -1
  double* gamma;           // n_inlet: dJ/dU_b of this step (inlet basis index)
};

hipError_t launch_adjoint(int Q, int precision, const AdjArgs& a, hipStream_t st);
// Marks in flag (npad bytes, cleared by the caller) each slot in entries[0, n)
// and every pull source of it: the state values a reverse step reads outside
// the moments tape.
hipError_t launch_spec_flags(int Q, int precision, const ColArgs& c, const int32_t* entries, int32_t n,
                             uint8_t* flag, hipStream_t st);
hipError_t launch_adjoint_boundary(int Q, int precision, const AdjArgs& a, hipStream_t st);
// out[t * n_modes + m] = sum_b gamma[t * n_inlet + b] * basis[m * n_inlet + b]
hipError_t launch_coef_gradient(const double* gamma, const double* basis, int32_t T, int32_t n_inlet,
                                int32_t n_modes, double* out, hipStream_t st);

}  // namespace lbm
