// Task and matrix records of the batched solve's task graph, shared by the
// device kernel (batched_dag.hip), the plan (batched_solve.cpp) and the
// scheduler (dag_schedule.h).  Plain C++: no HIP types.
#pragma once

namespace hc {
namespace dag {

constexpr int kDagTile = 64;

enum TaskType : int { kTaskDiag = 0, kTaskPanel = 1, kTaskTrtri = 2, kTaskLauum = 3, kTaskSolve = 4 };

struct DagTask {
  int m, type, i, j;  // matrix, task type, tile row, tile column
};

struct DagMat {
  int n, T, cbase, pad;  // order, tiles per side, first counter, unused
  long offA, offB, offP; // element offsets: A / Ainv, B / X / R, padded workspaces
};

}  // namespace dag
}  // namespace hc
