// aggregate.hpp -- greedy aggregation of a node graph for the multilevel preconditioner.
//
// Plain C++ (no device code): the multilevel set-up calls it on the host, and the sanitizer
// tests build it on its own (a ThreadSanitizer run of the two-thread form against the
// sequential one).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace cstfem_hip {

// Greedy aggregation of the graph (ptr, col) of n nodes, in node order: a node whose
// neighbours are all free becomes a root and takes its whole neighbourhood; leftovers join
// the aggregate of their first aggregated neighbour (column order); nodes with none start
// their own, numbered in node order after all others.  agg[i] = aggregate of node i; returns
// the number of aggregates.
int aggregate(const int* ptr, const int* col, int n, std::vector<int>& agg);

// The same aggregation with the leftover pass on a second thread, given the graph's
// bandwidth (|i - j| <= bandwidth for every edge); the result equals aggregate()'s exactly.
// Graphs under 100,000 nodes, or a bandwidth outside [0, n], take the sequential pass.
// `col` may still be arriving: with `need`, the root pass calls need(m) before it reads
// col[m - 1] beyond what an earlier call made available, and need returns how many leading
// entries of col are available from then on (at least m).  ptr must be complete.
int aggregate_pipelined(const int* ptr, const int* col, int n, int bandwidth,
                        std::vector<int>& agg,
                        const std::function<int64_t(int64_t)>& need = nullptr);

}  // namespace cstfem_hip
