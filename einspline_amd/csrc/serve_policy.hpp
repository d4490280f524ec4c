// serve_policy.hpp -- the serving runtime's admission decision, as a pure
// function of the tables' states (no GPU, no clock), so that it can be
// tested and used for planning on its own (einspline_server_plan_admission).
//
// A request needs table `want` resident.  If it is not, and the budget has
// no room for it, some resident tables must go.  Evicting a table costs its
// re-upload when it is needed again, so the victims are the set of tables
// with the fewest bytes that frees enough room (ties: fewer with queued
// requests, then least recently used).  A table with work in flight (or
// being uploaded, or already taken into the batch being formed) cannot be
// freed now; when the cheapest set holds such a table the decision is to
// wait for it ("drain": it gets no new work until it is idle) rather than
// to evict a larger idle table.  In the serving workload (48^3 x 256 FP32,
// 64^3 x 512 BF16, 96^3 x 1024 FP32 under 4.3 GB) this keeps the 4 GB table
// resident and swaps the two small ones.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace esk {

struct AdmitTable {
  size_t bytes;        // device bytes when resident
  int state;           // 0 absent, 1 resident, 2 being uploaded
  int busy;            // work in flight or taken into the batch being formed
  int pending;         // queued requests
  uint64_t last_use;   // logical clock of the last batch that used it
};

enum { ADMIT_RESIDENT = 0, ADMIT_UPLOAD = 1, ADMIT_DEFER = 2, ADMIT_NEVER = -1 };

// Subsets are enumerated exactly up to this many candidate tables; past it
// the candidates are taken greedily, idle before busy, least recently used
// first.
constexpr int kAdmitExact = 16;

inline int plan_admission(const std::vector<AdmitTable> &t, int want, size_t used, size_t budget,
                          std::vector<int> &victims, std::vector<int> &drain) {
  victims.clear();
  drain.clear();
  const AdmitTable &w = t[want];
  if (w.state == 1) return ADMIT_RESIDENT;
  if (w.state == 2) return ADMIT_DEFER;
  if (w.bytes > budget) return ADMIT_NEVER;
  if (used <= budget && w.bytes <= budget - used) return ADMIT_UPLOAD;
  const size_t shortfall = used + w.bytes - budget;
  std::vector<int> cand;
  for (int i = 0; i < (int)t.size(); ++i)
    if (i != want && t[i].state != 0) cand.push_back(i);
  const int n = (int)cand.size();
  std::vector<int> best;
  if (n <= kAdmitExact) {
    // exact: fewest bytes; ties: fewer tables with queued requests, then the
    // oldest most recent use among the set's members
    bool found = false;
    size_t best_bytes = 0;
    int best_pend = 0;
    uint64_t best_age = 0;
    for (uint32_t m = 1; m < (1u << n); ++m) {
      size_t b = 0;
      int pend = 0;
      uint64_t age = 0;
      for (int k = 0; k < n; ++k)
        if (m & (1u << k)) {
          const AdmitTable &c = t[cand[k]];
          b += c.bytes;
          pend += c.pending > 0;
          age = age > c.last_use ? age : c.last_use;
        }
      if (b < shortfall) continue;
      if (!found || b < best_bytes || (b == best_bytes && (pend < best_pend || (pend == best_pend && age < best_age)))) {
        found = true;
        best_bytes = b;
        best_pend = pend;
        best_age = age;
        best.clear();
        for (int k = 0; k < n; ++k)
          if (m & (1u << k)) best.push_back(cand[k]);
      }
    }
    if (!found) return ADMIT_DEFER;      // cannot happen while bytes <= budget and `used` counts only tables
  } else {
    std::vector<int> order = cand;
    for (size_t a = 1; a < order.size(); ++a)
      for (size_t b = a; b > 0; --b) {
        const AdmitTable &x = t[order[b - 1]], &y = t[order[b]];
        const bool xbusy = x.busy || x.state == 2, ybusy = y.busy || y.state == 2;
        if (xbusy > ybusy || (xbusy == ybusy && x.last_use > y.last_use)) {
          const int s = order[b - 1];
          order[b - 1] = order[b];
          order[b] = s;
        } else {
          break;
        }
      }
    size_t b = 0;
    for (int i : order) {
      if (b >= shortfall) break;
      best.push_back(i);
      b += t[i].bytes;
    }
    if (b < shortfall) return ADMIT_DEFER;
  }
  for (int i : best)
    if (t[i].busy || t[i].state == 2) drain.push_back(i);
  if (!drain.empty()) return ADMIT_DEFER;
  victims = best;
  return ADMIT_UPLOAD;
}

}  // namespace esk
