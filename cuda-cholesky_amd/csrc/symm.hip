// R = beta C + alpha A X for a symmetric FP64 matrix A stored as one triangle
// and a few right-hand sides X (n x nrhs, 16 columns per slice): the
// residual and the operator of the mixed-precision solver.
//
// Bound by the matrix cores and HBM together (A is read once; 16 columns make
// 32 multiply-adds per element of A, and the FP64 matrix cores of a CU consume
// 16 bytes of A per clock at that rate), so the stored triangle is read ONCE:
// every 64 x 64 tile T of the lower form (T(r, c) = A(r, c), or A(c, r) for
// upper storage) serves both
//   U: rows r  += T X_c         (the stored triangle)
//   V: rows c  += T^T X_r       (its mirror image; strict part on the diagonal)
// One workgroup per 512 x 512 region of the lower form: waves 0-3 form U (16
// rows of the tile each, one accumulator per row tile), waves 4-7 form V (16
// columns each, one accumulator per column tile of the region); both write
// their sums once as partial sums and a second pass adds a row block's
// partials in a fixed order (deterministic, no atomics).  Partials are 1/16
// of the region's bytes.
//
// Pipeline: the region's X rows (the V operand, 64 KiB) sit in LDS for the
// whole region; each tile and its X columns (the U operand) go global ->
// registers -> LDS, two tiles ahead in two register stages (measured on
// MI355X: four stages are no faster), so every product
// reads LDS only and the only waits on global loads come after a tile's
// products (the load counter completes in order: an operand load issued
// after a prefetch would otherwise wait for the prefetch).  Row tiles are the
// outer (rolled) loop and the 8 column tiles of a row the inner (unrolled)
// one, so the V accumulators are indexed at compile time.  Tiles of a
// diagonal region above its diagonal are neither loaded nor multiplied.
// Products use v_mfma_f64_16x16x4_f64.
#include "internal.h"

namespace hc {
namespace {

constexpr int RG = 512;               // region order
constexpr int TB = 64;                // tile order
constexpr int NTB = RG / TB;          // 8 tiles per region side
constexpr int LDT = TB + 2;           // LDS column stride of a tile (see the products below)
constexpr int LDX = 16 + 2;           // LDS row stride of a tile's X rows
constexpr int SY_THREADS = 512;       // 8 waves
constexpr int TV = TB * TB / 2 / SY_THREADS;  // 16-byte tile loads per thread (4)

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

// One prefetch stage: a thread's share of a tile and of its X columns.
struct Stage {
  v2d a[TV];
  v2d x;
};

template <bool LOWER>
__global__ void __launch_bounds__(SY_THREADS)
symm_region_kernel(int n, int nreg, const double *A, long lda, int vec, const double *Xc, long npad,
                   double *part) {
  __shared__ double Xr[RG * 16];        // X rows R0 .. R0 + 511 (V operand)
  __shared__ double Ts[2][TB * LDT];    // tiles, column-major
  __shared__ double Xt[2][TB * LDX];    // X rows of a tile's columns (U operand)
  int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int y = blockIdx.y;
  int a, b;
  region_of(blockIdx.x, nreg, a, b);
  const bool diag = a == b;
  const long R0 = (long)a * RG, C0 = (long)b * RG;
  const double *xc = Xc + (long)y * npad * 16;

  // X rows of the region into LDS (rows beyond n are zero in the compact copy)
#pragma unroll
  for (int u = 0; u < RG * 16 / 2 / SY_THREADS; u++) {
    const int e = tid + u * SY_THREADS, row = e >> 3, pr = 2 * (e & 7);
    *reinterpret_cast<v2d *>(Xr + row * 16 + pr) = *reinterpret_cast<const v2d *>(xc + (R0 + row) * 16 + pr);
  }

  // tile t = 8 p + q: row tile p, column tile q; above the diagonal of a
  // diagonal region (q > p) it is skipped
  auto skip = [&](int t) { return diag && (t & (NTB - 1)) > (t >> 3); };
  auto load = [&](Stage &st, int t) __attribute__((always_inline)) {
    if (t >= NTB * NTB || skip(t)) return;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // addresses formed here, not hoisted and kept live across the loop
    const int p = t >> 3, q = t & (NTB - 1);
    const long r0 = R0 + (long)p * TB, c0 = C0 + (long)q * TB;
    st.x = *reinterpret_cast<const v2d *>(xc + (c0 + (tid >> 3)) * 16 + 2 * (tid & 7));
    if (vec && r0 + TB <= n && c0 + TB <= n) {  // interior tile: plain 16-byte loads
      const double *base = LOWER ? A + (c0 + (tid >> 5)) * lda + r0 + 2 * (tid & 31)
                                 : A + (r0 + (tid >> 5)) * lda + c0 + 2 * (tid & 31);
#pragma unroll
      for (int u = 0; u < TV; u++) st.a[u] = *reinterpret_cast<const v2d *>(base + 16L * u * lda);
      return;
    }
#pragma unroll
    for (int u = 0; u < TV; u++) {
      const int v = tid + u * SY_THREADS;      // vector index 0 .. 2047
      const int i2 = 2 * (v & 31), j = v >> 5;  // pair along the contiguous index, other index
      // lower: (r, c) = (i2, j) and (i2 + 1, j); upper: (r, c) = (j, i2) and (j, i2 + 1)
      const long gr0 = LOWER ? r0 + i2 : r0 + j, gc0 = LOWER ? c0 + j : c0 + i2;
      const long gr1 = LOWER ? gr0 + 1 : gr0, gc1 = LOWER ? gc0 : gc0 + 1;
      const bool ok0 = gr0 < n && gc0 < n, ok1 = gr1 < n && gc1 < n;
      auto at = [&](long r, long c) {  // memory element of lower-form (r, c)
        const long rr = min(r, (long)n - 1), cc = min(c, (long)n - 1);
        return LOWER ? A[cc * lda + rr] : A[rr * lda + cc];
      };
      const double e0 = at(gr0, gc0), e1 = at(gr1, gc1);
      st.a[u] = (v2d){ok0 ? e0 : 0.0, ok1 ? e1 : 0.0};
    }
  };
  auto store = [&](const Stage &st, int t) __attribute__((always_inline)) {
    if (t >= NTB * NTB || skip(t)) return;
    double *T = Ts[t & 1];
    *reinterpret_cast<v2d *>(Xt[t & 1] + (tid >> 3) * LDX + 2 * (tid & 7)) = st.x;
#pragma unroll
    for (int u = 0; u < TV; u++) {
      const int v = tid + u * SY_THREADS, i2 = 2 * (v & 31), j = v >> 5;
      if (LOWER) {
        *reinterpret_cast<v2d *>(T + j * LDT + i2) = st.a[u];
      } else {
        T[i2 * LDT + j] = st.a[u][0];
        T[(i2 + 1) * LDT + j] = st.a[u][1];
      }
    }
  };

  // partial sums of this region: U rows (R0 ..) and V rows (C0 ..), 512 x 16 each
  double *pu = part + ((long)y * nreg * (nreg + 1) + 2L * blockIdx.x) * RG * 16;
  double *pv = pu + RG * 16;

  const bool uwave = wave < 4;
  const int w4 = wave & 3;
  v4d V[NTB];
#pragma unroll
  for (int q = 0; q < NTB; q++) V[q] = (v4d){0.0, 0.0, 0.0, 0.0};

  // NST register stages: tiles t + 1 .. t + NST - 1 in flight while tile t
  // is multiplied (t mod NST = q mod NST: the stage is known at compile time)
  constexpr int NST = 2;
  Stage st[NST];
#pragma unroll
  for (int k = 0; k < NST; k++) load(st[k], k);
  store(st[0], 0);
  __syncthreads();
#pragma unroll 1
  for (int p = 0; p < NTB; p++) {
    v4d U = (v4d){0.0, 0.0, 0.0, 0.0};
    // V operand of every tile of row tile p: X rows 64 p + 4 u + lane / 16
    double xv[TB / 4];
    if (!uwave) {
#pragma unroll
      for (int u = 0; u < TB / 4; u++) xv[u] = Xr[(p * TB + 4 * u + lg) * 16 + lc];
    }
#pragma unroll
    for (int q = 0; q < NTB; q++) {
      const int t = NTB * p + q;
      Stage &cur = st[q % NST], &nxt = st[(q + 1) % NST];
      // tile t + NST into the stage tile t came through (stored last iteration)
      load(cur, t + NST);
      asm volatile("" ::: "memory");  // keep the prefetch ahead of the products
      if (!(diag && q > p)) {
        const double *T = Ts[t & 1];
        const bool md = diag && q == p;  // diagonal tile: lower part only
        // LDS banks: a half-wave (lane groups g = lane / 16 of 0, 1 or 2, 3)
        // reads 32 doubles at once.  V lanes differ in the column (stride
        // LDT = 66 doubles: 16 distinct even bank pairs) and g in the row.
        // U lanes differ in the row, so the two lane groups of a half-wave
        // take columns 8 apart (8 LDT doubles: the other 32 banks) and their
        // X rows are LDX = 18 doubles apart for the same reason.
        if (uwave) {  // U rows 16 w4 + lc: sum over the tile's 64 columns
          const int r = 16 * w4 + lc;
          const double *xt = Xt[t & 1];
#pragma unroll
          for (int u = 0; u < TB / 4; u++) {
            const int c = (u & 7) + 8 * lg + 32 * (u >> 3);  // a permutation of the 64 columns
            double av = T[c * LDT + r];
            if (md && r < c) av = 0.0;
            U = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xt[c * LDX + lc], U, 0, 0, 0);
          }
        } else {      // V column 16 w4 + lc: sum over the tile's 64 rows (strict part on the diagonal)
          const int c = 16 * w4 + lc;
#pragma unroll
          for (int u = 0; u < TB / 4; u++) {
            const int r = 4 * u + lg;
            double av = T[c * LDT + r];
            if (md && r <= c) av = 0.0;
            V[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xv[u], V[q], 0, 0, 0);
          }
        }
      }
      // tile t + 1 (loaded one iteration ago) into the buffer tile t - 1 used,
      // which every wave finished reading before the last barrier
      store(nxt, t + 1);
      __syncthreads();
    }
    // U of row tile p: wave w4 rows 64 p + 16 w4 + (lane / 16 + 4 r), column lane % 16
    if (uwave) {
#pragma unroll
      for (int r = 0; r < 4; r++) pu[(p * TB + 16 * w4 + lg + 4 * r) * 16 + lc] = U[r];
    }
  }
  if (!uwave) {
#pragma unroll
    for (int q = 0; q < NTB; q++)
#pragma unroll
      for (int r = 0; r < 4; r++) pv[(q * TB + 16 * w4 + lg + 4 * r) * 16 + lc] = V[q][r];
  }
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
