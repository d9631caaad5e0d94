// Task graph and list schedule of the batched solve (host side, plain C++).
//
// add_matrix() appends the tasks of one matrix of order n (T = ceil(n / 64)
// tiles per side) with their dependencies: DIAG(j) and PANEL(i, j) after
// column j - 1 of the factor is complete (PANEL also after DIAG(j)),
// TRTRI(i, j) after DIAG(i) and TRTRI(i - 1, j), LAUUM after every TRTRI and
// DIAG(T - 1), SOLVE after every LAUUM.  Zero-cost pseudo nodes stand for
// "column j complete", "inverse of L complete" and "inverse complete" so
// that the edge count stays linear in the task count.
//
// schedule() simulates a list schedule on `workers` identical workers (the
// kernel's persistent workgroups) with the ready task of longest remaining
// path taken first, and returns the real tasks in the order they start:
// a topological order in which the tasks on long chains come early.
//
// check_order() replays an order on ONE worker with the device's wait rules
// (counter thresholds, not graph edges) and returns the index of the first
// task whose wait could never be satisfied by the tasks before it, or -1.
// A kernel whose workgroups take tasks in list order and only wait on tasks
// earlier in the list cannot deadlock; the plan refuses an order that fails
// this check.
#pragma once

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "dag_types.h"

namespace hc {
namespace dag {

struct TaskGraph {
  std::vector<DagTask> task;   // pseudo nodes have type -1
  std::vector<double> cost;    // estimated microseconds
  std::vector<std::pair<int, int>> edges;
  int add(int m, int type, int i, int j, double c) {
    task.push_back(DagTask{m, type, i, j});
    cost.push_back(c);
    return (int)task.size() - 1;
  }
};

// Estimated durations (microseconds): products at kRate flop/us per
// workgroup, a fixed cost per task, the diagonal factor + inverse, and the
// solve as about five sweeps over two n x n FP64 matrices at kBw bytes/us.
struct CostModel {
  double rate = 8.0e4, task = 2.0, diag = 20.0, bw = 1.0e5;
  double prod(double k) const { return 2.0 * kDagTile * kDagTile * k / rate; }
};

inline void add_matrix(TaskGraph &g, int m, int n, const CostModel &cm = CostModel()) {
  constexpr int TB = kDagTile;
  const int T = (n + TB - 1) / TB;
  if (T == 0) {
    g.add(m, kTaskSolve, 0, 0, cm.task);
    return;
  }
  std::vector<int> diag(T), colp(T);
  std::vector<std::vector<int>> trtri(T);
  for (int j = 0; j < T; j++) {
    diag[j] = g.add(m, kTaskDiag, j, j, cm.task + cm.prod(TB * j) + cm.diag);
    colp[j] = g.add(m, -1, 0, j, 0.0);
    g.edges.push_back({diag[j], colp[j]});
    if (j > 0) g.edges.push_back({colp[j - 1], diag[j]});
    for (int i = j + 1; i < T; i++) {
      const int t = g.add(m, kTaskPanel, i, j, cm.task + cm.prod(TB * j) + cm.prod(TB));
      g.edges.push_back({diag[j], t});
      g.edges.push_back({t, colp[j]});
      if (j > 0) g.edges.push_back({colp[j - 1], t});
    }
  }
  const int tt = g.add(m, -1, 0, 0, 0.0);
  g.edges.push_back({diag[T - 1], tt});
  for (int i = 1; i < T; i++) {
    trtri[i].assign(i, -1);
    for (int j = 0; j < i; j++) {
      const int t = g.add(m, kTaskTrtri, i, j, cm.task + cm.prod(TB * (i - j)) + cm.prod(TB));
      trtri[i][j] = t;
      g.edges.push_back({diag[i], t});
      if (i - 1 > j) g.edges.push_back({trtri[i - 1][j], t});
      g.edges.push_back({t, tt});
    }
  }
  const int lt = g.add(m, -1, 0, 0, 0.0);
  for (int i = 0; i < T; i++)
    for (int j = 0; j <= i; j++) {
      const int t = g.add(m, kTaskLauum, i, j, cm.task + cm.prod(TB * (T - i)));
      g.edges.push_back({tt, t});
      g.edges.push_back({t, lt});
    }
  const int sv = g.add(m, kTaskSolve, 0, 0, cm.task + 5.0 * 2.0 * 8.0 * (double)n * n / cm.bw);
  g.edges.push_back({lt, sv});
}

inline std::vector<DagTask> schedule(const TaskGraph &g, int workers) {
  const int N = (int)g.task.size();
  std::vector<int> head(N + 1, 0), adj(g.edges.size()), indeg(N, 0);
  for (const auto &e : g.edges) {
    head[e.first + 1]++;
    indeg[e.second]++;
  }
  for (int u = 0; u < N; u++) head[u + 1] += head[u];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (const auto &e : g.edges) adj[fill[e.first]++] = e.second;
  }
  // topological order (Kahn), then bottom levels in reverse
  std::vector<int> topo;
  topo.reserve(N);
  {
    std::vector<int> deg(indeg);
    for (int u = 0; u < N; u++)
      if (deg[u] == 0) topo.push_back(u);
    for (size_t q = 0; q < topo.size(); q++)
      for (int e = head[topo[q]]; e < head[topo[q] + 1]; e++)
        if (--deg[adj[e]] == 0) topo.push_back(adj[e]);
  }
  std::vector<double> bl(N, 0.0);
  for (int q = (int)topo.size() - 1; q >= 0; q--) {
    const int u = topo[q];
    double best = 0.0;
    for (int e = head[u]; e < head[u + 1]; e++) best = std::max(best, bl[adj[e]]);
    bl[u] = g.cost[u] + best;
  }
  // the list schedule
  std::priority_queue<std::pair<double, int>> ready;  // (bottom level, -id)
  std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>,
                      std::greater<std::pair<double, int>>> running;  // (finish, id)
  std::vector<int> deg(indeg);
  for (int u = 0; u < N; u++)
    if (deg[u] == 0) ready.push({bl[u], -u});
  std::vector<DagTask> order;
  order.reserve(N);
  double now = 0.0;
  int idle = std::max(1, workers);
  auto complete = [&](int u) {
    for (int e = head[u]; e < head[u + 1]; e++)
      if (--deg[adj[e]] == 0) ready.push({bl[adj[e]], -adj[e]});
  };
  for (;;) {
    while (idle > 0 && !ready.empty()) {
      const int u = -ready.top().second;
      ready.pop();
      if (g.task[u].type < 0) {  // pseudo node: completes at once
        complete(u);
        continue;
      }
      order.push_back(g.task[u]);
      running.push({now + g.cost[u], u});
      idle--;
    }
    if (running.empty()) break;
    const auto f = running.top();
    running.pop();
    now = f.first;
    idle++;
    complete(f.second);
  }
  return order;
}

// Sequential replay of `order` with the kernel's wait rules; Ts[m] is the
// tile count of matrix m.  Returns the first task that would wait forever,
// -1 if none, or -2 if a task is missing or repeated.
inline long check_order(const std::vector<DagTask> &order, const std::vector<int> &Ts) {
  const size_t M = Ts.size();
  std::vector<std::vector<int>> diag(M), col(M), trow(M);
  std::vector<int> ttot(M, 0), ltot(M, 0), solved(M, 0);
  std::vector<long> expect(M, 0), seen(M, 0);
  for (size_t m = 0; m < M; m++) {
    const long T = Ts[m];
    diag[m].assign(T, 0);
    col[m].assign(T, 0);
    trow[m].assign(T, 0);
    expect[m] = T + T * (T - 1) / 2 + T * (T - 1) / 2 + T * (T + 1) / 2 + 1;
  }
  for (size_t t = 0; t < order.size(); t++) {
    const DagTask &k = order[t];
    if (k.m < 0 || (size_t)k.m >= M) return -2;
    const int m = k.m, T = Ts[m], i = k.i, j = k.j;
    seen[m]++;
    switch (k.type) {
      case kTaskDiag:
        if (j < 0 || j >= T || i != j || diag[m][j]) return -2;
        if (j > 0 && col[m][j - 1] < T - j + 1) return (long)t;
        diag[m][j] = 1;
        col[m][j]++;
        break;
      case kTaskPanel:
        if (j < 0 || i <= j || i >= T) return -2;
        if (j > 0 && col[m][j - 1] < T - j + 1) return (long)t;
        if (diag[m][j] < 1) return (long)t;
        col[m][j]++;
        break;
      case kTaskTrtri:
        if (j < 0 || i <= j || i >= T) return -2;
        if (diag[m][i] < 1) return (long)t;
        if (i - j - 1 > 0 && trow[m][j] < i - j - 1) return (long)t;
        // the kernel's rows of a column finish top-down: row i is next
        if (trow[m][j] != i - j - 1) return -2;
        trow[m][j]++;
        ttot[m]++;
        break;
      case kTaskLauum:
        if (j < 0 || i < j || i >= T) return -2;
        if (ttot[m] < T * (T - 1) / 2 || diag[m][T - 1] < 1) return (long)t;
        ltot[m]++;
        break;
      case kTaskSolve:
        if (solved[m]) return -2;
        if (ltot[m] < T * (T + 1) / 2) return (long)t;
        solved[m] = 1;
        break;
      default:
        return -2;
    }
  }
  for (size_t m = 0; m < M; m++)
    if (seen[m] != expect[m] || !solved[m]) return -2;
  return -1;
}

}  // namespace dag
}  // namespace hc
