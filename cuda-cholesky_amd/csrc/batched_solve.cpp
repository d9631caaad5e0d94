// Batched SPD solve + log-determinant + inverse of matrices of any size (no
// counterpart in the reference; its routines take one matrix per call, the
// per-matrix loop being lapack/dpotrf.c:248-286 followed by dpotri.c:25-46).
//
// A plan fixes the sizes and offsets of a batch: it builds the task graph of
// 64 x 64 tiles of every matrix (factor, inverse, solve; batched_dag.hip),
// orders it by a list-scheduling simulation over the kernel's persistent
// workgroups (critical path first, so that the largest matrices, whose
// dependency chains are longest, start early) and allocates the padded
// workspaces.  A run is ONE persistent kernel (plus the clearing of its
// counters), captured in a HIP graph on first use and replayed with a single
// graph launch while the buffers stay the same.
//
// Per matrix i (n_i x n_i, column-major, ld n_i, at element offset offA[i];
// right-hand sides and solutions n_i x nrhs, ld n_i, at offB[i]):
//   * the FP64 Cholesky factor and its inverse X = inv(L) (workspace);
//   * A_i^-1 = X^T X, written to Ainv (triangle uplo) when Ainv is given;
//   * X_i = A_i^-1 B_i refined in FP64 against A_i (residuals with the
//     original matrix) until LAPACK DSPOSV's stopping test holds, at most
//     30 steps: iter[i] = steps taken, -31 without convergence;
//   * logdet[i] = sum of log L_jj^2 (NaN on failure);
//   * info[i] = 0, or the order of the leading minor that is not positive
//     definite, or -1 when a dependency wait was abandoned.
// A is never written.
#include "api_util.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "batched_dag.h"
#include "dag_schedule.h"
#include "cucholesky.h"

using hc::bad_arg;
using hc::dag::DagMat;
using hc::dag::DagParams;
using hc::dag::DagTask;

struct __cubatchedsolveplan_st {
  int device;
  char uplo;
  int count, nrhs, grid, ntasks;
  long nints;                // ticket, error word and counters
  DagTask *d_tasks;
  DagMat *d_mats;
  int *d_ints;
  double *d_logpart;
  double *F, *Xw, *Af, *R;
  // graph of the last run
  hipStream_t capture;
  hipGraph_t graph;
  hipGraphExec_t exec;
  CUdeviceptr g[7];
};

namespace {

constexpr int TB = hc::dag::kDagTile;

void release_graph(__cubatchedsolveplan_st *p) {
  if (p->exec != nullptr) (void)hipGraphExecDestroy(p->exec);
  if (p->graph != nullptr) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

CUresult plan_release(__cubatchedsolveplan_st *p) {
  CUresult first = CUDA_SUCCESS;
  auto keep = [&](hipError_t e) {
    if (e != hipSuccess && first == CUDA_SUCCESS) first = hc::to_cu(e);
  };
  if (p->capture != nullptr) keep(hipStreamSynchronize(p->capture));
  release_graph(p);
  if (p->capture != nullptr) keep(hipStreamDestroy(p->capture));
  void *bufs[] = {p->d_tasks, p->d_mats, p->d_ints, p->d_logpart, p->F, p->Xw, p->Af, p->R};
  for (void *b : bufs)
    if (b != nullptr) keep(hipFree(b));
  std::free(p);
  return first;
}

}  // namespace

extern "C" {

CUresult cuDsposvxBatchedPlanCreate(CULAPACKhandle h, CBlasUplo uplo, size_t count, const size_t *n,
                                    const size_t *offA, size_t nrhs, const size_t *offB,
                                    CUbatchedsolveplan *plan) {
  const char *name = "cuDsposvxBatchedPlanCreate";
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (count > (size_t)0x7fffffff) return bad_arg(name, 3);
  if (count > 0 && n == nullptr) return bad_arg(name, 4);
  if (count > 0 && offA == nullptr) return bad_arg(name, 5);
  if (nrhs > (size_t)0x7fffffff) return bad_arg(name, 6);
  if (count > 0 && nrhs > 0 && offB == nullptr) return bad_arg(name, 7);
  if (plan == nullptr) return bad_arg(name, 8);
  for (size_t i = 0; i < count; i++)
    if (n[i] > (size_t)0x100000) return bad_arg(name, 4);
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  int cus = 0;
  HC_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const int grid = std::max(1, cus * hc::dag::dag_blocks_per_cu(uplo == CBlasUpper));

  // matrices: tiles, counters, padded workspace offsets
  std::vector<DagMat> mats(count);
  long cb = 0, op = 0;
  size_t rlen = 0;
  hc::dag::TaskGraph gr;
  std::vector<int> Ts(count);
  for (size_t i = 0; i < count; i++) {
    const int ni = (int)n[i], T = (ni + TB - 1) / TB;
    mats[i] = DagMat{ni, T, (int)cb, 0, (long)offA[i], nrhs > 0 ? (long)offB[i] : 0L, op};
    cb += 3L * T + 3;
    op += (long)TB * T * TB * T;
    if (nrhs > 0) rlen = std::max(rlen, (size_t)offB[i] + (size_t)ni * nrhs);
    hc::dag::add_matrix(gr, (int)i, ni);
    Ts[i] = T;
  }
  if (cb + 2 > 0x7fffffffL) return bad_arg(name, 3);
  const std::vector<DagTask> order = hc::dag::schedule(gr, grid);
  // the kernel only waits on tasks earlier in the list: refuse any order
  // that would make it wait on a later one
  if (hc::dag::check_order(order, Ts) != -1) {
    hc::report("task order check", name, __FILE__, __LINE__, CUDA_ERROR_UNKNOWN);
    return CUDA_ERROR_UNKNOWN;
  }

  auto *p = static_cast<__cubatchedsolveplan_st *>(std::calloc(1, sizeof(__cubatchedsolveplan_st)));
  if (p == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  p->device = h->device;
  p->uplo = (char)uplo;
  p->count = (int)count;
  p->nrhs = (int)nrhs;
  p->grid = grid;
  p->ntasks = (int)order.size();
  p->nints = 2 + cb;
  hipError_t e = hipSuccess;
  auto alloc = [&](void **ptr, size_t bytes) {
    if (e == hipSuccess && bytes > 0) e = hipMalloc(ptr, bytes);
  };
  alloc(reinterpret_cast<void **>(&p->d_tasks), order.size() * sizeof(DagTask));
  alloc(reinterpret_cast<void **>(&p->d_mats), count * sizeof(DagMat));
  alloc(reinterpret_cast<void **>(&p->d_ints), (size_t)p->nints * sizeof(int));
  alloc(reinterpret_cast<void **>(&p->d_logpart), std::max<long>(cb, 1) * sizeof(double));
  alloc(reinterpret_cast<void **>(&p->F), (size_t)op * sizeof(double));
  alloc(reinterpret_cast<void **>(&p->Xw), (size_t)op * sizeof(double));
  alloc(reinterpret_cast<void **>(&p->Af), (size_t)op * sizeof(double));
  alloc(reinterpret_cast<void **>(&p->R), rlen * sizeof(double));
  if (e == hipSuccess && !order.empty())
    e = hipMemcpy(p->d_tasks, order.data(), order.size() * sizeof(DagTask), hipMemcpyHostToDevice);
  if (e == hipSuccess && count > 0)
    e = hipMemcpy(p->d_mats, mats.data(), count * sizeof(DagMat), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)plan_release(p);
    HC_CHECK(e);
  }
  *plan = p;
  return CUDA_SUCCESS;
}

CUresult cuDsposvxBatchedRun(CUbatchedsolveplan p, CUdeviceptr A, CUdeviceptr B, CUdeviceptr X,
                             CUdeviceptr Ainv, CUdeviceptr iter, CUdeviceptr logdet, CUdeviceptr info,
                             CUstream stream) {
  const char *name = "cuDsposvxBatchedRun";
  if (p == nullptr) return bad_arg(name, 1);
  if (p->count == 0) return CUDA_SUCCESS;
  if (A == 0) return bad_arg(name, 2);
  if (p->nrhs > 0 && B == 0) return bad_arg(name, 3);
  if (p->nrhs > 0 && X == 0) return bad_arg(name, 4);
  if (iter == 0) return bad_arg(name, 6);
  if (logdet == 0) return bad_arg(name, 7);
  if (info == 0) return bad_arg(name, 8);
  hc::DeviceGuard g(p->device);
  HC_CHECK(g.status());
  const CUdeviceptr key[7] = {A, B, X, Ainv, iter, logdet, info};
  bool same = p->exec != nullptr;
  for (int k = 0; k < 7 && same; k++) same = p->g[k] == key[k];
  if (!same) {
    release_graph(p);
    DagParams d{};
    d.tasks = p->d_tasks;
    d.ntasks = p->ntasks;
    d.mats = p->d_mats;
    d.tick = p->d_ints;
    d.err = p->d_ints + 1;
    d.cnt = p->d_ints + 2;
    d.A = hc::dptr<const double>(A);
    d.B = hc::dptr<const double>(B);
    d.X = hc::dptr<double>(X);
    d.Ainv = Ainv != 0 ? hc::dptr<double>(Ainv) : nullptr;
    d.F = p->F;
    d.Xw = p->Xw;
    d.Af = p->Af;
    d.R = p->R;
    d.logpart = p->d_logpart;
    d.iter = hc::dptr<int>(iter);
    d.logdet = hc::dptr<double>(logdet);
    d.info = hc::dptr<int>(info);
    d.nrhs = p->nrhs;
    HC_CHECK(hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal));
    const hipError_t e = hc::dag::dag_run(p->capture, p->uplo == 'U', d, p->nints, p->grid);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(p->capture, &graph);
    hipError_t e3 = e == hipSuccess ? e2 : e;
    if (e3 == hipSuccess) e3 = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      if (graph != nullptr) (void)hipGraphDestroy(graph);
      p->exec = nullptr;
      HC_CHECK(e3);
    }
    p->graph = graph;
    for (int k = 0; k < 7; k++) p->g[k] = key[k];
  }
  HC_CHECK(hipGraphLaunch(p->exec, stream));
  return CUDA_SUCCESS;
}

CUresult cuDsposvxBatchedPlanDestroy(CUbatchedsolveplan p) {
  if (p == nullptr) return CUDA_SUCCESS;
  hc::DeviceGuard g(p->device);
  HC_CHECK(g.status());
  HC_CHECK(hipDeviceSynchronize());
  return plan_release(p);
}

CUresult cuDsposvxBatched(CULAPACKhandle h, CBlasUplo uplo, size_t count, const size_t *n,
                          const size_t *offA, CUdeviceptr A, size_t nrhs, const size_t *offB, CUdeviceptr B,
                          CUdeviceptr X, CUdeviceptr Ainv, long *iter, double *logdet, long *info) {
  const char *name = "cuDsposvxBatched";
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (count > 0 && n == nullptr) return bad_arg(name, 4);
  if (count > 0 && offA == nullptr) return bad_arg(name, 5);
  if (count > 0 && A == 0) return bad_arg(name, 6);
  if (count > 0 && nrhs > 0 && offB == nullptr) return bad_arg(name, 8);
  if (count > 0 && nrhs > 0 && (B == 0 || X == 0)) return bad_arg(name, B == 0 ? 9 : 10);
  if (count > 0 && (iter == nullptr || logdet == nullptr || info == nullptr))
    return bad_arg(name, iter == nullptr ? 12 : (logdet == nullptr ? 13 : 14));
  if (count == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  CUbatchedsolveplan p = nullptr;
  CUresult r = cuDsposvxBatchedPlanCreate(h, uplo, count, n, offA, nrhs, offB, &p);
  if (r != CUDA_SUCCESS) return r;
  int *d_out = nullptr;  // iter (int), info (int), logdet (double), err copy
  const size_t ib = (2 * count * sizeof(int) + 7) / 8 * 8;
  hipError_t e = hipMalloc(&d_out, ib + count * sizeof(double));
  std::vector<int> it(count), inf(count);
  std::vector<double> ld(count);
  int err = 0;
  if (e == hipSuccess) {
    int *d_iter = d_out, *d_info = d_out + count;
    double *d_log = reinterpret_cast<double *>(reinterpret_cast<char *>(d_out) + ib);
    r = cuDsposvxBatchedRun(p, A, B, X, Ainv, reinterpret_cast<CUdeviceptr>(d_iter),
                            reinterpret_cast<CUdeviceptr>(d_log), reinterpret_cast<CUdeviceptr>(d_info),
                            h->stream);
    if (r == CUDA_SUCCESS) e = hipMemcpyAsync(it.data(), d_iter, count * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (r == CUDA_SUCCESS && e == hipSuccess)
      e = hipMemcpyAsync(inf.data(), d_info, count * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (r == CUDA_SUCCESS && e == hipSuccess)
      e = hipMemcpyAsync(ld.data(), d_log, count * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (r == CUDA_SUCCESS && e == hipSuccess)
      e = hipMemcpyAsync(&err, p->d_ints + 1, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (r == CUDA_SUCCESS && e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_out);
  }
  const CUresult rd = cuDsposvxBatchedPlanDestroy(p);
  if (r != CUDA_SUCCESS) return r;
  HC_CHECK(e);
  if (rd != CUDA_SUCCESS) return rd;
  for (size_t i = 0; i < count; i++) {
    iter[i] = it[i];
    info[i] = inf[i];
    logdet[i] = ld[i];
  }
  if (err != 0) {  // a dependency wait was abandoned: the results are not valid
    hc::report("dependency wait abandoned", name, __FILE__, __LINE__, CUDA_ERROR_LAUNCH_FAILED);
    return CUDA_ERROR_LAUNCH_FAILED;
  }
  return CUDA_SUCCESS;
}

}  // extern "C"
