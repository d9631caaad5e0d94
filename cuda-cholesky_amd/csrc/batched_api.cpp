// Batched Cholesky + log-determinant (+ inverse) of many small SPD matrices
// of uneven sizes (no counterpart in the reference; its routines take one
// matrix per call).  A plan fixes the sizes, offsets and schedule once; each
// run is one persistent kernel, captured in a HIP graph on first use and
// replayed with a single graph launch.
#include "api_util.h"

#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "cucholesky.h"

using hc::bad_arg;

struct __cubatchedplan_st {
  int device;
  char uplo;
  char precision;  // 'D' (FP64 storage) or 'S' (FP32 storage)
  int count;
  int grid_inverse;        // persistent grid of the factor + inverse kernel
  int reg_count[3], reg_grid[3];  // register-resident factor kernels: matrices and grid per size class
  int *d_n;
  long *d_off;
  int *d_order;
  int *d_cnt;  // reg_count on the device
  // graph for the last (A, logdet, info, inverse) run
  hipStream_t capture;
  hipGraph_t graph;
  hipGraphExec_t exec;
  CUdeviceptr gA, gLog, gInfo;
  bool gInverse;
};

namespace {

hipError_t launch(const __cubatchedplan_st *p, hipStream_t s, CUdeviceptr A, CUdeviceptr logdet,
                  CUdeviceptr info, bool inverse) {
  double *ld = reinterpret_cast<double *>(logdet);
  int *inf = reinterpret_cast<int *>(info);
  if (!inverse) {
    if (p->precision == 'D')
      return hc::potrf_batched_reg<double>(s, p->uplo, p->d_cnt, p->reg_count, p->reg_grid, p->d_n, p->d_off,
                                           p->d_order, reinterpret_cast<double *>(A), ld, inf);
    return hc::potrf_batched_reg<float>(s, p->uplo, p->d_cnt, p->reg_count, p->reg_grid, p->d_n, p->d_off,
                                        p->d_order, reinterpret_cast<float *>(A), ld, inf);
  }
  if (p->precision == 'D')
    return hc::potrf_batched<double>(s, p->uplo, p->count, p->d_n, p->d_off, p->d_order,
                                     reinterpret_cast<double *>(A), ld, inf, p->grid_inverse, true);
  return hc::potrf_batched<float>(s, p->uplo, p->count, p->d_n, p->d_off, p->d_order,
                                  reinterpret_cast<float *>(A), ld, inf, p->grid_inverse, true);
}

void release_graph(__cubatchedplan_st *p) {
  if (p->exec != nullptr) (void)hipGraphExecDestroy(p->exec);
  if (p->graph != nullptr) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

// Frees everything a (possibly partly built) plan holds; every member is
// null or valid.
CUresult plan_release(__cubatchedplan_st *p) {
  CUresult first = CUDA_SUCCESS;
  auto keep = [&](hipError_t e) {
    if (e != hipSuccess && first == CUDA_SUCCESS) first = hc::to_cu(e);
  };
  if (p->capture != nullptr) keep(hipStreamSynchronize(p->capture));
  release_graph(p);
  if (p->capture != nullptr) keep(hipStreamDestroy(p->capture));
  if (p->d_n != nullptr) keep(hipFree(p->d_n));
  if (p->d_off != nullptr) keep(hipFree(p->d_off));
  if (p->d_order != nullptr) keep(hipFree(p->d_order));
  if (p->d_cnt != nullptr) keep(hipFree(p->d_cnt));
  std::free(p);
  return first;
}

// One graph launch of the batch; the graph is re-captured when the buffers
// or the operation change.
CUresult run(const char *name, CUbatchedplan p, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
             CUstream stream, bool inverse) {
  if (p == nullptr) return bad_arg(name, 1);
  if (p->count > 0 && A == 0) return bad_arg(name, 2);
  if (p->count > 0 && info == 0) return bad_arg(name, 4);
  if (p->count == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(p->device);
  HC_CHECK(g.status());
  if (p->exec == nullptr || p->gA != A || p->gLog != logdet || p->gInfo != info ||
      p->gInverse != inverse) {
    release_graph(p);
    HC_CHECK(hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal));
    const hipError_t e = launch(p, p->capture, A, logdet, info, inverse);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(p->capture, &graph);
    hipError_t e3 = (e == hipSuccess) ? e2 : e;
    if (e3 == hipSuccess) e3 = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {  // no half-built graph is kept
      if (graph != nullptr) (void)hipGraphDestroy(graph);
      p->exec = nullptr;
      HC_CHECK(e3);
    }
    p->graph = graph;
    p->gA = A;
    p->gLog = logdet;
    p->gInfo = info;
    p->gInverse = inverse;
  }
  HC_CHECK(hipGraphLaunch(p->exec, stream));
  return CUDA_SUCCESS;
}

}  // namespace

extern "C" {

CUresult cuPotrfBatchedPlanCreate(CULAPACKhandle h, CBlasUplo uplo, char precision, size_t count,
                                  const size_t *n, const size_t *offsets, CUbatchedplan *plan) {
  const char *name = "cuPotrfBatchedPlanCreate";
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (precision != 'D' && precision != 'S') return bad_arg(name, 3);
  if (count > (size_t)0x7fffffff) return bad_arg(name, 4);
  if (count > 0 && n == nullptr) return bad_arg(name, 5);
  if (count > 0 && offsets == nullptr) return bad_arg(name, 6);
  if (plan == nullptr) return bad_arg(name, 7);
  for (size_t i = 0; i < count; i++)
    if (n[i] > (size_t)hc::kBatchedMaxN) return bad_arg(name, 5);
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  int cus = 0;
  HC_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  auto *p = static_cast<__cubatchedplan_st *>(std::calloc(1, sizeof(__cubatchedplan_st)));
  if (p == nullptr) return CUDA_ERROR_OUT_OF_MEMORY;
  p->device = h->device;
  p->uplo = (char)uplo;
  p->precision = precision;
  p->count = (int)count;
  // longest first: the round-robin deal over the persistent grid then
  // balances uneven sizes (work grows as n^3)
  std::vector<int> order(count), nv(count);
  std::vector<long> off(count);
  for (size_t i = 0; i < count; i++) {
    nv[i] = (int)n[i];
    off[i] = (long)offsets[i];
  }
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nv[a] > nv[b]; });
  // persistent grid: at most every CU filled to the kernel's occupancy (a
  // workgroup that waits for a slot would start its share late); fewer per CU
  // when the autotuner measured that faster ($CUCHOLESKY_BATCHED_PER_CU
  // overrides both)
  const char *env = std::getenv("CUCHOLESKY_BATCHED_PER_CU");
  const int forced = env != nullptr ? std::atoi(env) : 0;
  auto per_cu = [&](bool inverse) {
    const int occ = precision == 'D' ? hc::potrf_batched_blocks_per_cu<double>((char)uplo, inverse)
                                     : hc::potrf_batched_blocks_per_cu<float>((char)uplo, inverse);
    const int want = forced > 0 ? forced : hc::tuned_batched_per_cu(precision, inverse);
    return want > 0 ? std::min(want, occ) : occ;
  };
  p->grid_inverse = std::max(1, std::min<int>((int)count, cus * per_cu(true)));
  // register-resident factor kernels: every CU filled to their occupancy
  for (size_t i = 0; i < count; i++) p->reg_count[hc::batched_reg_class(nv[i])]++;
  for (int c = 0; c < 3; c++) {
    const int occ = precision == 'D' ? hc::potrf_batched_reg_blocks_per_cu<double>((char)uplo, c)
                                     : hc::potrf_batched_reg_blocks_per_cu<float>((char)uplo, c);
    p->reg_grid[c] = cus * (forced > 0 ? std::min(forced, occ) : occ);
  }
  hipError_t e = hipMalloc(&p->d_cnt, 3 * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(p->d_cnt, p->reg_count, 3 * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && count > 0) {
    e = hipMalloc(&p->d_n, count * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_off, count * sizeof(long));
    if (e == hipSuccess) e = hipMalloc(&p->d_order, count * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(p->d_n, nv.data(), count * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_off, off.data(), count * sizeof(long), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_order, order.data(), count * sizeof(int), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)plan_release(p);
    HC_CHECK(e);
  }
  *plan = p;
  return CUDA_SUCCESS;
}

CUresult cuPotrfBatchedRun(CUbatchedplan p, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
                           CUstream stream) {
  return run("cuPotrfBatchedRun", p, A, logdet, info, stream, false);
}

CUresult cuPotriBatchedRun(CUbatchedplan p, CUdeviceptr A, CUdeviceptr logdet, CUdeviceptr info,
                           CUstream stream) {
  return run("cuPotriBatchedRun", p, A, logdet, info, stream, true);
}

CUresult cuPotrfBatchedPlanDestroy(CUbatchedplan p) {
  if (p == nullptr) return CUDA_SUCCESS;
  hc::DeviceGuard g(p->device);
  HC_CHECK(g.status());
  HC_CHECK(hipDeviceSynchronize());
  return plan_release(p);
}

}  // extern "C"
