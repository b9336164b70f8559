// Binomial checkpointing schedule (Griewank's "revolve") for reversing a
// run of time steps with a bounded number of stored states.
//
// reverse(l, n, s): steps l+1 .. l+n are reversed in decreasing order with
// the state l held in a buffer and s free state buffers: the first stored
// state is placed m steps ahead, the tail is reversed with s - 1 buffers and
// the head with s. m minimises the forward steps
//   cost(n, s) = min_m [ m + cost(n - m, s - 1) + cost(m, s) ],
// whose optimum is Griewank's r n - C(s + r, s + 1) for the least r with
// C(s + r, s) >= n. s >= n - 1 stores every state (n - 1 forward steps);
// s = 0 recomputes every state from l.
#pragma once

#include <cstdint>
#include <vector>

namespace lbm {

constexpr int32_t REV_ADVANCE = 0;  // state `from` (buffer `buf_from`) -> state `to` into `buf_to`
constexpr int32_t REV_REVERSE = 1;  // reverse step `to` = from + 1 using state `from` in `buf_from`
constexpr int32_t REV_BASE = -1;    // buffer holding the state l
constexpr int32_t REV_SCRATCH = -2; // extra buffer used when no slot is free

struct RevAction {
  int32_t op;
  int32_t from;
  int32_t to;
  int32_t buf_from;
  int32_t buf_to;
};

// Actions reversing steps l+1 .. l+n with `slots` buffers numbered 0..slots-1.
std::vector<RevAction> revolve_schedule(int32_t l, int32_t n, int32_t slots);

}  // namespace lbm
