// The two-thread aggregation (aggregate_pipelined: the leftover pass on a second thread,
// synchronised by a published progress counter) against the sequential pass, on node graphs
// above the 100,000-node threshold where the second thread is used; and the same with the
// column array still arriving: a producer thread copies it in pieces, each published under
// a mutex, and the aggregation's `need` callback waits for the pieces it reaches (the way the
// device download's events are waited for).  Built with -fsanitize=thread by the sanitizer
// test: ThreadSanitizer reports any access pair the progress counter, the mutex or the
// callback does not order.  Exit status 0 when every graph matches.
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "aggregate.hpp"

struct Graph {
  std::vector<int> ptr, col;
  int n = 0, bandwidth = 0;
};

// node graph of an nx x ny grid of nodes with the neighbours of a CST mesh whose cells are
// split along one diagonal (left, right, up, down and the two diagonal neighbours), itself
// included as the mesh pattern has it; `perm` renumbers the nodes (new index of node i)
Graph grid_graph(int nx, int ny, const std::vector<int>& perm) {
  Graph g;
  g.n = nx * ny;
  std::vector<std::vector<int>> adj(g.n);
  const int di[7] = {0, -1, 1, 0, 0, -1, 1}, dj[7] = {0, 0, 0, -1, 1, -1, 1};
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < 7; ++k) {
        const int a = i + di[k], b = j + dj[k];
        if (a < 0 || a >= nx || b < 0 || b >= ny) continue;
        adj[perm[j * nx + i]].push_back(perm[b * nx + a]);
      }
  g.ptr.assign(1, 0);
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    for (int c : a) g.bandwidth = std::max(g.bandwidth, std::abs(c - (int)(&a - adj.data())));
    g.col.insert(g.col.end(), a.begin(), a.end());
    g.ptr.push_back((int)g.col.size());
  }
  return g;
}

// aggregate_pipelined while another thread fills col in `pieces` pieces
int pipelined_arriving(const Graph& g, int pieces, std::vector<int>& out) {
  const int64_t m = (int64_t)g.col.size();
  std::vector<int> col(m, -1);
  std::mutex mu;
  std::condition_variable cv;
  int64_t arrived = 0;  // leading entries of col published
  std::thread producer([&]() {
    for (int p = 0; p < pieces; ++p) {
      const int64_t lo = m * p / pieces, hi = m * (p + 1) / pieces;
      std::copy(g.col.begin() + lo, g.col.begin() + hi, col.begin() + lo);
      std::this_thread::yield();
      {
        std::lock_guard<std::mutex> lk(mu);
        arrived = hi;
      }
      cv.notify_all();
    }
  });
  auto need = [&](int64_t want) -> int64_t {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&]() { return arrived >= want; });
    return arrived;
  };
  const int na = cstfem_hip::aggregate_pipelined(g.ptr.data(), col.data(), g.n, g.bandwidth, out,
                                                 need);
  producer.join();
  return na;
}

int check(const char* name, const Graph& g) {
  std::vector<int> a, b, c;
  const int na = cstfem_hip::aggregate(g.ptr.data(), g.col.data(), g.n, a);
  const int nb = cstfem_hip::aggregate_pipelined(g.ptr.data(), g.col.data(), g.n, g.bandwidth, b);
  const int nc = pipelined_arriving(g, 8, c);
  const bool ok = na == nb && a == b && na == nc && a == c;
  std::printf("%s: %d nodes, bandwidth %d, %d aggregates: %s\n", name, g.n, g.bandwidth, na,
              ok ? "identical" : "DIFFERENT");
  return ok ? 0 : 1;
}

int main() {
  int fails = 0;
  const int nx = 420, ny = 400;
  std::vector<int> id(nx * ny);
  std::iota(id.begin(), id.end(), 0);
  fails += check("grid 420x400", grid_graph(nx, ny, id));
  // renumbered within windows of 2000 nodes: a larger bandwidth, leftovers in other orders
  std::vector<int> p = id;
  std::mt19937 rng(7);
  for (int w = 0; w < nx * ny; w += 2000)
    std::shuffle(p.begin() + w, p.begin() + std::min(nx * ny, w + 2000), rng);
  fails += check("grid 420x400, windowed renumbering", grid_graph(nx, ny, p));
  // isolated nodes between the rows' blocks: nodes with no aggregated neighbour (phase 3)
  Graph g = grid_graph(nx, ny, id);
  Graph h;
  h.n = g.n + g.n / 50;
  h.ptr.assign(1, 0);
  h.bandwidth = 0;
  std::vector<int> map(g.n);
  for (int i = 0, k = 0; i < g.n; ++i, ++k) {
    if (i % 50 == 0) ++k;  // an isolated node before every 50th node
    map[i] = k;
  }
  std::vector<std::vector<int>> adj(h.n);
  for (int i = 0; i < g.n; ++i)
    for (int q = g.ptr[i]; q < g.ptr[i + 1]; ++q) adj[map[i]].push_back(map[g.col[q]]);
  for (int i = 0; i < h.n; ++i) {
    if (adj[i].empty()) adj[i].push_back(i);
    for (int c : adj[i]) h.bandwidth = std::max(h.bandwidth, std::abs(c - i));
    h.col.insert(h.col.end(), adj[i].begin(), adj[i].end());
    h.ptr.push_back((int)h.col.size());
  }
  fails += check("grid with isolated nodes", h);
  return fails;
}
