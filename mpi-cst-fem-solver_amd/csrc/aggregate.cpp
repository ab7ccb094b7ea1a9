// aggregate.cpp -- greedy aggregation of a node graph (declarations and contract in
// aggregate.hpp).
#include "aggregate.hpp"

#include <atomic>
#include <cstdint>
#include <exception>
#include <thread>

namespace cstfem_hip {

int aggregate(const int* ptr, const int* col, int n, std::vector<int>& agg) {
  agg.assign(n, -1);
  int na = 0;
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int q = ptr[i]; q < ptr[i + 1] && free_nb; ++q) free_nb = agg[col[q]] < 0;
    if (!free_nb) continue;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) agg[col[q]] = na;
    agg[i] = na;
    ++na;
  }
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (agg[col[q]] >= 0) {
        agg[i] = agg[col[q]];
        break;
      }
  }
  for (int i = 0; i < n; ++i)
    if (agg[i] < 0) agg[i] = na++;
  return na;
}

// The same aggregation with the leftover pass (phase 2) on a second thread.  Given the
// graph's bandwidth B (|i - j| <= B for every edge), phase 1 settles node k once it has
// passed node k + B (only roots within B of k can claim it), so the leftover pass may treat
// node i (reading i and its neighbours, all <= i + B) once phase 1 has passed i + 2B; phase 1
// never reads or writes below its own position minus B again.  The leftover pass runs in node
// order as in aggregate(), and the new aggregates of phase 3 are numbered in node order after
// both, so the result is aggregate()'s exactly.
int aggregate_pipelined(const int* ptr, const int* col, int n, int bandwidth,
                        std::vector<int>& agg,
                        const std::function<int64_t(int64_t)>& need) {
  if (n < 100000 || bandwidth < 0 || bandwidth > n) {
    if (need) need(n > 0 ? ptr[n] : 0);  // the sequential passes read all of col
    return aggregate(ptr, col, n, agg);
  }
  agg.assign(n, -1);
  std::atomic<int> progress{0};  // phase 1 has finished the nodes below this
  std::vector<int> fail;         // leftovers with no aggregated neighbour (phase 3)
  int* a = agg.data();
  const int64_t lag = 2 * (int64_t)bandwidth + 1;
  std::exception_ptr error;
  std::atomic<bool> abort{false};  // phase 1 failed (its `need` threw): stop, read nothing more
  std::thread leftovers([&]() {
    try {
      int seen = 0;
      for (int i = 0; i < n; ++i) {
        while ((int64_t)seen < (int64_t)i + lag && seen < n && !abort.load(std::memory_order_acquire))
          seen = progress.load(std::memory_order_acquire);
        if (abort.load(std::memory_order_acquire)) break;
        if (a[i] >= 0) continue;
        int g = -1;
        for (int q = ptr[i]; q < ptr[i + 1]; ++q)
          if (a[col[q]] >= 0) {
            g = a[col[q]];
            break;
          }
        if (g >= 0)
          a[i] = g;
        else
          fail.push_back(i);
      }
    } catch (...) {
      error = std::current_exception();
    }
  });
  int na = 0;
  // leading entries of col known to be there; the published progress i promises the
  // leftover pass (which reads the rows below it) that rows 0 .. i-1 have arrived
  int64_t have = need ? 0 : INT64_MAX;
  try {
    for (int i = 0; i < n; ++i) {
      if ((i & 1023) == 0) {
        if (ptr[i] > have) have = need(ptr[i]);
        progress.store(i, std::memory_order_release);
      }
      if (a[i] >= 0) continue;
      if (ptr[i + 1] > have) have = need(ptr[i + 1]);
      bool free_nb = true;
      for (int q = ptr[i]; q < ptr[i + 1] && free_nb; ++q) free_nb = a[col[q]] < 0;
      if (!free_nb) continue;
      for (int q = ptr[i]; q < ptr[i + 1]; ++q) a[col[q]] = na;
      a[i] = na;
      ++na;
    }
    if (ptr[n] > have) have = need(ptr[n]);
  } catch (...) {
    abort.store(true, std::memory_order_release);
    leftovers.join();
    throw;
  }
  progress.store(n, std::memory_order_release);
  leftovers.join();
  if (error) std::rethrow_exception(error);
  for (int i : fail)
    if (a[i] < 0) a[i] = na++;
  return na;
}

}  // namespace cstfem_hip
