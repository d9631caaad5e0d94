// Host check of the batched solve's task scheduler (dag_schedule.h), built
// by tests/test_dag_schedule.py with g++.
//   dag_schedule_check <seed> <count> <nmax> <workers>
// prints one line: "<tasks> <check> <mut1> <mut2> <mut3>" where check is
// check_order() of the schedule (-1 = every wait satisfiable) and mut1..3 are
// check_order() of three corrupted copies: a DIAG moved before the column it
// waits for, a TRTRI moved before its predecessor in the column, and a task
// removed (each must be rejected: >= 0 or -2).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "dag_schedule.h"

using namespace hc::dag;

int main(int argc, char **argv) {
  if (argc < 5) return 2;
  const unsigned seed = (unsigned)atoi(argv[1]);
  const int count = atoi(argv[2]), nmax = atoi(argv[3]), workers = atoi(argv[4]);
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> dist(0, nmax);
  TaskGraph g;
  std::vector<int> Ts(count);
  for (int m = 0; m < count; m++) {
    const int n = m == 0 ? nmax : dist(rng);  // at least one large matrix
    add_matrix(g, m, n);
    Ts[m] = (n + kDagTile - 1) / kDagTile;
  }
  const std::vector<DagTask> order = schedule(g, workers);
  const long check = check_order(order, Ts);
  // 1: the first DIAG(j > 0) of matrix 0 moved to the front
  std::vector<DagTask> o1 = order;
  for (size_t t = 0; t < o1.size(); t++)
    if (o1[t].m == 0 && o1[t].type == kTaskDiag && o1[t].j > 0) {
      const DagTask k = o1[t];
      o1.erase(o1.begin() + t);
      o1.insert(o1.begin(), k);
      break;
    }
  // 2: the last TRTRI of matrix 0 swapped with the TRTRI above it in its column
  std::vector<DagTask> o2 = order;
  long last = -1;
  for (size_t t = 0; t < o2.size(); t++)
    if (o2[t].m == 0 && o2[t].type == kTaskTrtri && o2[t].i - o2[t].j >= 2) last = (long)t;
  if (last >= 0)
    for (long t = 0; t < last; t++)
      if (o2[t].m == 0 && o2[t].type == kTaskTrtri && o2[t].j == o2[last].j && o2[t].i == o2[last].i - 1) {
        std::swap(o2[t], o2[last]);
        break;
      }
  // 3: a LAUUM task dropped
  std::vector<DagTask> o3 = order;
  for (size_t t = 0; t < o3.size(); t++)
    if (o3[t].type == kTaskLauum) {
      o3.erase(o3.begin() + t);
      break;
    }
  std::printf("%zu %ld %ld %ld %ld\n", order.size(), check, check_order(o1, Ts), check_order(o2, Ts),
              check_order(o3, Ts));
  return 0;
}
