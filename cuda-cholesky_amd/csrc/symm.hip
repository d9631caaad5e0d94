// R = beta C + alpha A X for a symmetric FP64 matrix A stored as one triangle
// and a few right-hand sides X (n x nrhs, 16 columns per slice): the
// residual and the operator of the mixed-precision solver.
//
// Memory-bound (A is read once; 16 columns make 32 flops per element of A),
// so the stored triangle is read ONCE: every 128 x 64 tile T of the lower
// form (T(r, c) = A(r, c), or A(c, r) for upper storage) serves both
//   U: rows r  += T X_c         (the stored triangle)
//   V: rows c  += T^T X_r       (its mirror image; strict part on the diagonal)
// One workgroup per 512 x 512 region of the lower form keeps its U sums
// (512 rows) and V sums (512 columns) in registers and writes them once as
// partial sums; a second pass adds a row block's partials in a fixed order
// (deterministic, no atomics).  Partials are 1/16 of the region's bytes.
//
// Tiles go global -> registers -> LDS (double-buffered, the next tile's loads
// in flight during the current tile's products); X comes from a compact copy
// (16 consecutive values per row) so that its operand loads are contiguous.
// Products use v_mfma_f64_16x16x4_f64.
#include "internal.h"

namespace hc {
namespace {

constexpr int RG = 512;               // region order
constexpr int TR = 128, TC = 64;      // tile rows x columns
constexpr int LDT = TR + 2;           // LDS column stride of a tile
constexpr int SY_THREADS = 512;       // 8 waves
constexpr int NTP = RG / TR, NTQ = RG / TC;  // 4 row tiles, 8 column tiles per region

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// compact copy: Xc[y][row][16] (rows padded with zeros up to npad)
__global__ void compact_kernel(long n, long npad, long nrhs, const double *X, long ldx, double *Xc) {
  const int y = blockIdx.y;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < npad * 16; e += (long)gridDim.x * blockDim.x) {
    const long r = e >> 4, c = 16L * y + (e & 15);
    Xc[(long)y * npad * 16 + e] = (r < n && c < nrhs) ? X[c * ldx + r] : 0.0;
  }
}

// region t of the lower triangle of RG-blocks, column-major over (a >= b)
__device__ __forceinline__ void region_of(int t, int nreg, int &a, int &b) {
  int bb = 0, rem = t;
  // column b holds nreg - b regions (a = b .. nreg - 1)
  while (rem >= nreg - bb) {
    rem -= nreg - bb;
    bb++;
  }
  b = bb;
  a = bb + rem;
}

template <bool LOWER>
__global__ void __launch_bounds__(SY_THREADS)
symm_region_kernel(int n, int nreg, const double *A, long lda, int vec, const double *Xc, long npad,
                   double *part) {
  __shared__ double Ts[2][TC * LDT];
  __shared__ double Vx[TC * 16];  // V of one column tile from waves 4..7
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int y = blockIdx.y;
  int a, b;
  region_of(blockIdx.x, nreg, a, b);
  const bool diag = a == b;
  const long R0 = (long)a * RG, C0 = (long)b * RG;
  const double *xc = Xc + (long)y * npad * 16;

  // global -> registers: thread owns 16 elements of the 128 x 64 tile
  auto load_tile = [&](v2d (&reg)[8], int p, int q) __attribute__((always_inline)) {
    const long r0 = R0 + (long)p * TR, c0 = C0 + (long)q * TC;
    if (vec && r0 + TR <= n && c0 + TC <= n) {  // interior tile: plain 16-byte loads
      const double *base = LOWER ? A + (c0 + wave) * lda + r0 + 2 * lane
                                 : A + (r0 + (tid >> 5)) * lda + c0 + 2 * (lane & 31);
#pragma unroll
      for (int u = 0; u < 8; u++)
        reg[u] = *reinterpret_cast<const v2d *>(base + (LOWER ? 8L * u * lda : 16L * u * lda));
      return;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int v = tid + u * SY_THREADS;  // vector index 0 .. 4095
      int r, c;
      if (LOWER) { r = 2 * (v & 63); c = v >> 6; }   // pairs along rows (contiguous)
      else { c = 2 * (v & 31); r = v >> 5; }         // pairs along columns (contiguous)
      const long gr = r0 + r, gc = c0 + c;
      v2d x;
      if (LOWER) {
        const long cc = min(gc, (long)n - 1);
        if (vec && gr + 1 < n && gc < n) {
          x = *reinterpret_cast<const v2d *>(A + cc * lda + gr);
        } else {
          const double e0 = A[cc * lda + min(gr, (long)n - 1)];
          const double e1 = A[cc * lda + min(gr + 1, (long)n - 1)];
          x = (v2d){e0, e1};
        }
        x[0] = (gr < n && gc < n) ? x[0] : 0.0;
        x[1] = (gr + 1 < n && gc < n) ? x[1] : 0.0;
      } else {
        const long rr = min(gr, (long)n - 1);
        if (vec && gc + 1 < n && gr < n) {
          x = *reinterpret_cast<const v2d *>(A + rr * lda + gc);
        } else {
          const double e0 = A[rr * lda + min(gc, (long)n - 1)];
          const double e1 = A[rr * lda + min(gc + 1, (long)n - 1)];
          x = (v2d){e0, e1};
        }
        x[0] = (gr < n && gc < n) ? x[0] : 0.0;
        x[1] = (gr < n && gc + 1 < n) ? x[1] : 0.0;
      }
      reg[u] = x;
    }
  };
  auto store_tile = [&](const v2d (&reg)[8], int buf) __attribute__((always_inline)) {
    double *T = Ts[buf];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int v = tid + u * SY_THREADS;
      if (LOWER) {
        const int r = 2 * (v & 63), c = v >> 6;
        *reinterpret_cast<v2d *>(T + c * LDT + r) = reg[u];
      } else {
        const int c = 2 * (v & 31), r = v >> 5;
        T[c * LDT + r] = reg[u][0];
        T[(c + 1) * LDT + r] = reg[u][1];
      }
    }
  };

  // partial sums of this region: U rows (R0 ..) and V rows (C0 ..), 512 x 16 each
  double *pu = part + ((long)y * nreg * (nreg + 1) + 2L * blockIdx.x) * RG * 16;
  double *pv = pu + RG * 16;

  v4d U[NTP], V;
#pragma unroll
  for (int p = 0; p < NTP; p++) U[p] = (v4d){0.0, 0.0, 0.0, 0.0};
  V = (v4d){0.0, 0.0, 0.0, 0.0};
  const int vr = 64 * (wave >> 2);      // V: this wave's half of the tile rows
  const int vc = 16 * (wave & 3) + lc;  // V: this lane's tile column

  // products of the tile in buffer buf (row tile p, column tile q)
  auto tile_products = [&](int buf, int p, int q, v4d &Up) __attribute__((always_inline)) {
    const double *T = Ts[buf];
    const long gr0 = R0 + (long)p * TR, gc0 = C0 + (long)q * TC;
    {  // U rows 16 wave + lc: sum over the tile's 64 columns
      const int r = 16 * wave + lc;
      v4d acc = Up;
#pragma unroll
      for (int u = 0; u < TC / 4; u++) {
        const int c = 4 * u + lg;
        double av = T[c * LDT + r];
        if (diag && gr0 + r < gc0 + c) av = 0.0;  // above the diagonal
        const double bv = xc[(gc0 + c) * 16 + lc];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
      }
      Up = acc;
    }
    {  // V column vc: sum over this wave's 64 tile rows of T(r, vc) X(r)
      v4d acc = V;
#pragma unroll
      for (int u = 0; u < 64 / 4; u++) {
        const int r = vr + 4 * u + lg;
        double av = T[vc * LDT + r];
        if (diag && gr0 + r <= gc0 + vc) av = 0.0;  // strict lower part only
        const double bv = xc[(gr0 + r) * 16 + lc];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
      }
      V = acc;
    }
  };

  // The 32 tiles of the region, q outer and p inner: the next tile's loads
  // are in flight during this tile's products.  On a diagonal region the
  // masks zero what lies above the diagonal.
  v2d ra[8];
  load_tile(ra, 0, 0);
  store_tile(ra, 0);
  __syncthreads();
#pragma unroll 1
  for (int t = 0; t < NTP * NTQ; t++) {
    const int p = t % NTP, q = t / NTP;
    const bool more = t + 1 < NTP * NTQ;
    if (more) load_tile(ra, (t + 1) % NTP, (t + 1) / NTP);
    tile_products(t & 1, p, q, U[p]);
    // column tile q complete: waves 4..7 hand their half to waves 0..3
    if (p == NTP - 1 && wave >= 4) {
#pragma unroll
      for (int r = 0; r < 4; r++) Vx[(16 * (wave & 3) + lg + 4 * r) * 16 + lc] = V[r];
    }
    // the other buffer was last read in the previous iteration, which every
    // wave finished before that iteration's barrier
    if (more) store_tile(ra, (t + 1) & 1);
    __syncthreads();
    if (p == NTP - 1) {
      if (wave < 4) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = 16 * wave + lg + 4 * r;
          pv[(q * TC + row) * 16 + lc] = V[r] + Vx[row * 16 + lc];
        }
      }
      V = (v4d){0.0, 0.0, 0.0, 0.0};
    }
  }

  // U: wave w, row-tile p: rows 128 p + 16 w + (lane / 16 + 4 r), column lane % 16
#pragma unroll
  for (int p = 0; p < NTP; p++)
#pragma unroll
    for (int r = 0; r < 4; r++) pu[(p * TR + 16 * wave + lg + 4 * r) * 16 + lc] = U[p][r];
}

// R(rows of block a) = beta C + alpha (sum_b U(a, b) + sum_{a' >= a} V(a', a));
// workgroup (a, chunk): 256 of the block's 512 x 16 values, one per thread,
// partials added in a fixed order.
constexpr int RD_THREADS = 256, RD_CHUNKS = RG * 16 / RD_THREADS;
__global__ void __launch_bounds__(RD_THREADS)
symm_reduce_kernel(int n, int nrhs, int nreg, const double *part, double alpha, double beta,
                   const double *C, long ldc, double *R, long ldr) {
  const int a = blockIdx.x / RD_CHUNKS, y = blockIdx.y;
  const int e = (blockIdx.x % RD_CHUNKS) * RD_THREADS + threadIdx.x;
  const double *base = part + (long)y * nreg * (nreg + 1) * RG * 16 + e;
  auto slot = [&](int ra, int rb) {  // region (ra >= rb), column-major over the triangle
    return (long)rb * nreg - (long)rb * (rb - 1) / 2 + (ra - rb);
  };
  const int rl = e >> 4, cl = e & 15;
  const long r = (long)a * RG + rl;
  const int c = 16 * y + cl;
  if (r >= n || c >= nrhs) return;
  double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
  for (int bb = 0; bb <= a; bb++) s0 += base[2 * slot(a, bb) * RG * 16];          // U
#pragma unroll 4
  for (int aa = a; aa < nreg; aa++) s1 += base[(2 * slot(aa, a) + 1) * RG * 16];  // V
  const double cv = beta == 0.0 ? 0.0 : beta * C[(long)c * ldc + r];
  R[(long)c * ldr + r] = cv + alpha * (s0 + s1);
}

}  // namespace

size_t symm_tall_ws_bytes(long n, long nrhs) {
  const long nreg = (n + RG - 1) / RG, gy = (nrhs + 15) / 16, npad = nreg * RG;
  const size_t part = (size_t)gy * nreg * (nreg + 1) * RG * 16 * sizeof(double);
  const size_t xc = (size_t)gy * npad * 16 * sizeof(double);
  return (part + 255) / 256 * 256 + xc;
}

hipError_t symm_tall(hipStream_t s, char uplo, long n, long nrhs, double alpha, const double *A,
                     long lda, const double *X, long ldx, double beta, const double *C, long ldc,
                     double *R, long ldr, void *ws) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (n > (long)INT32_MAX / 2 || nrhs > 16L * 65535) return hipErrorInvalidValue;
  const long nreg = (n + RG - 1) / RG, gy = (nrhs + 15) / 16, npad = nreg * RG;
  double *part = static_cast<double *>(ws);
  const size_t part_bytes = ((size_t)gy * nreg * (nreg + 1) * RG * 16 * sizeof(double) + 255) / 256 * 256;
  double *Xc = reinterpret_cast<double *>(static_cast<char *>(ws) + part_bytes);
  hipLaunchKernelGGL(compact_kernel, dim3((unsigned)std::min<long>((npad * 16 + 255) / 256, 4096), (unsigned)gy),
                     dim3(256), 0, s, n, npad, nrhs, X, ldx, Xc);
  HC_TRY(hipGetLastError());
  const int vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0 && lda % 2 == 0) ? 1 : 0;
  const dim3 grid((unsigned)(nreg * (nreg + 1) / 2), (unsigned)gy);
  if (uplo == 'L')
    hipLaunchKernelGGL(symm_region_kernel<true>, grid, dim3(SY_THREADS), 0, s, (int)n, (int)nreg, A, lda,
                       vec, Xc, npad, part);
  else
    hipLaunchKernelGGL(symm_region_kernel<false>, grid, dim3(SY_THREADS), 0, s, (int)n, (int)nreg, A, lda,
                       vec, Xc, npad, part);
  HC_TRY(hipGetLastError());
  hipLaunchKernelGGL(symm_reduce_kernel, dim3((unsigned)(nreg * RD_CHUNKS), (unsigned)gy), dim3(RD_THREADS), 0, s, (int)n,
                     (int)nrhs, (int)nreg, part, alpha, beta, C, ldc, R, ldr);
  return hipGetLastError();
}

}  // namespace hc
