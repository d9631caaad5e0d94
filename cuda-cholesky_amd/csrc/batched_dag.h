// Task-graph kernel of the batched SPD solve + logdet + inverse
// (batched_dag.hip) and the plan that feeds it (batched_solve.cpp).
#pragma once

#include "internal.h"
#include "dag_types.h"

namespace hc {
namespace dag {

struct DagParams {
  const DagTask *tasks;
  int ntasks;
  const DagMat *mats;
  int *tick;        // ticket counter; the error word and the counters follow it
  int *err;
  int *cnt;
  const double *A, *B;
  double *X, *Ainv;  // Ainv may be null
  double *F, *Xw, *Af, *R;
  double *logpart;   // one per diagonal tile, at cbase + j
  int *iter;
  double *logdet;
  int *info;
  int nrhs;
};

// Resident workgroups per CU of the task kernel.
int dag_blocks_per_cu(bool upper);
// Clears the ncounters ints from p.tick (ticket, error word, counters) and
// runs the task kernel on `grid` persistent workgroups.
hipError_t dag_run(hipStream_t s, bool upper, const DagParams &p, long ncounters, int grid);

}  // namespace dag
}  // namespace hc
