#include "checkpoint.h"

#include <algorithm>
#include <climits>
#include <memory>

#include "../capi_internal.h"

namespace lbm {
namespace {

// C(s + t, s), in double (exact for the sizes used, saturating beyond).
double beta(int32_t s, int32_t t) {
  double r = 1.0;
  for (int32_t i = 1; i <= s; ++i) r = r * double(t + i) / double(i);
  return r;
}

// Minimal forward-step counts cost(n, s) = min_m [m + cost(n - m, s - 1) +
// cost(m, s)] and their best first split m, for n <= N, s <= S.
struct SplitTable {
  int32_t N = 0, S = 0;
  std::vector<int64_t> cost;
  std::vector<int32_t> split;
  int64_t& c(int32_t n, int32_t s) { return cost[static_cast<size_t>(n) * (S + 1) + s]; }
  int32_t& m(int32_t n, int32_t s) { return split[static_cast<size_t>(n) * (S + 1) + s]; }
  SplitTable(int32_t n_max, int32_t s_max) : N(n_max), S(s_max) {
    cost.assign(static_cast<size_t>(N + 1) * (S + 1), 0);
    split.assign(cost.size(), 1);
    for (int32_t n = 2; n <= N; ++n) {
      c(n, 0) = int64_t(n) * (n - 1) / 2;
      for (int32_t s = 1; s <= S; ++s) {
        int64_t best = INT64_MAX;
        int32_t arg = 1;
        for (int32_t k = 1; k < n; ++k) {
          const int64_t v = k + c(n - k, s - 1) + c(k, s);
          if (v < best) {
            best = v;
            arg = k;
          }
        }
        c(n, s) = best;
        m(n, s) = arg;
      }
    }
  }
};

// First split: from the table when present, else the binomial rule
// m = max(1, n - beta(s - 1, t)) with t the least sweep count reaching n.
int32_t first_split(int32_t n, int32_t s, SplitTable* tab) {
  if (s >= n - 1) return 1;
  if (tab) return tab->m(n, s);
  int32_t t = 1;
  while (beta(s, t) < double(n)) ++t;
  const double tail = std::min(beta(s - 1, t), double(n - 1));
  return std::max<int32_t>(1, n - static_cast<int32_t>(tail));
}

void plan(int32_t l, int32_t n, int32_t buf_l, std::vector<int32_t>& free_bufs, SplitTable* tab,
          std::vector<RevAction>& out) {
  if (n <= 0) return;
  if (n == 1) {
    out.push_back({REV_REVERSE, l, l + 1, buf_l, REV_SCRATCH});
    return;
  }
  const int32_t s = static_cast<int32_t>(free_bufs.size());
  if (s == 0) {
    for (int32_t k = n; k >= 1; --k) {
      if (k == 1) {
        out.push_back({REV_REVERSE, l, l + 1, buf_l, REV_SCRATCH});
      } else {
        out.push_back({REV_ADVANCE, l, l + k - 1, buf_l, REV_SCRATCH});
        out.push_back({REV_REVERSE, l + k - 1, l + k, REV_SCRATCH, REV_SCRATCH});
      }
    }
    return;
  }
  const int32_t m = first_split(n, s, tab);
  const int32_t buf_m = free_bufs.back();
  free_bufs.pop_back();
  out.push_back({REV_ADVANCE, l, l + m, buf_l, buf_m});
  plan(l + m, n - m, buf_m, free_bufs, tab, out);
  free_bufs.push_back(buf_m);
  plan(l, m, buf_l, free_bufs, tab, out);
}

}  // namespace

std::vector<RevAction> revolve_schedule(int32_t l, int32_t n, int32_t slots) {
  std::vector<int32_t> free_bufs;
  for (int32_t i = std::max<int32_t>(slots, 0) - 1; i >= 0; --i) free_bufs.push_back(i);
  std::vector<RevAction> out;
  const int32_t s = static_cast<int32_t>(free_bufs.size());
  // The exact table costs O(n^2 s); beyond that the binomial rule is used.
  std::unique_ptr<SplitTable> tab;
  if (s > 0 && s < n - 1 && double(n) * n * s <= 2e8) tab = std::make_unique<SplitTable>(n, s);
  plan(l, n, REV_BASE, free_bufs, tab.get(), out);
  return out;
}

}  // namespace lbm

extern "C" int lbm_checkpoint_schedule(int32_t nsteps, int32_t slots, int32_t* actions, int64_t capacity,
                                       int64_t* count) {
  if (nsteps < 0 || slots < 0 || !count) return lbm::fail(LBM_ERR_ARG, "invalid checkpoint schedule request");
  const auto plan = lbm::revolve_schedule(0, nsteps, slots);
  *count = static_cast<int64_t>(plan.size());
  if (actions) {
    if (capacity < *count) return lbm::fail(LBM_ERR_ARG, "action buffer too small");
    for (size_t i = 0; i < plan.size(); ++i) {
      actions[5 * i] = plan[i].op;
      actions[5 * i + 1] = plan[i].from;
      actions[5 * i + 2] = plan[i].to;
      actions[5 * i + 3] = plan[i].buf_from;
      actions[5 * i + 4] = plan[i].buf_to;
    }
  }
  return LBM_OK;
}
