// Public LAPACK entry points (include/lapack.h): xPOTRF, xTRTRI, xTRTI22,
// xLAUUM, xPOTRI and xLOGDET for double and double complex.
//
// Every step runs on the GPU.  The reference copies each diagonal block to the
// host, factors it there and copies it back (lapack/dpotrf.c:263-275); here a
// single-workgroup LDS kernel factors the block and also returns its inverse,
// so the panel solve becomes a matrix-core GEMM and the loop never waits for
// the host.  `info` lives in device memory; kernels issued after a failed
// pivot see it and do nothing, and one copy at the end returns it.
#include "api_util.h"

#include <cstdlib>

#include <algorithm>
#include <type_traits>

using hc::bad_arg;
using hc::dptr;
using hc::zd;

namespace {

// Handle workspace of at least `bytes`, grown (never shrunk) on demand.
hipError_t workspace(CULAPACKhandle h, size_t bytes, void **out) {
  if (bytes > h->ws_bytes) {
    if (h->ws != nullptr) {
      HC_TRY(hipStreamSynchronize(h->stream));
      HC_TRY(hipStreamSynchronize(h->side));
      HC_TRY(hipFree(h->ws));
      h->ws = nullptr;
      h->ws_bytes = 0;
    }
    bytes = (bytes + 0xFFFFF) & ~size_t(0xFFFFF);  // 1 MiB granules
    HC_TRY(hipMalloc(&h->ws, bytes));
    h->ws_bytes = bytes;
  }
  *out = h->ws;
  return hipSuccess;
}

template <typename T> constexpr int conj_op() { return hc::is_complex<T>::value ? hc::OP_C : hc::OP_T; }

// Outer panel width of the right-looking factorisation: the trailing update
// has inner dimension NB, i.e. NB / 8 flops per byte of the trailing matrix
// it streams (double), enough to stay matrix-core bound.
// Outer block size: the handle's override, else $CUCHOLESKY_POTRF_NB, else
// the measured table, else the built-in default.  prec: D, Z, S or B.
int prec_slot(char prec) { return prec == 'D' ? 0 : prec == 'Z' ? 1 : prec == 'S' ? 2 : 3; }

long potrf_nb(CULAPACKhandle h, char prec, long n) {
  const long o = h != nullptr ? h->nb_override[prec_slot(prec)] : 0;
  if (o > 0) return o;
  static const long env = [] {
    const char *e = std::getenv("CUCHOLESKY_POTRF_NB");
    return e ? std::atol(e) : 0L;
  }();
  if (env > 0) return env;
  const long t = hc::tuned_potrf_nb(prec, n);
  if (t > 0) return t;
  return prec == 'Z' ? 128L : 384L;
}

template <typename T> char prec_of(bool bf16) {
  if (hc::is_complex<T>::value) return 'Z';
  if (std::is_same<T, float>::value) return bf16 ? 'B' : 'S';
  return 'D';
}

// Panel factorisation of columns [j, j + jb) (lower; rows [j, j + jb) for
// upper), all rows below, in IB-column slices: diagonal block (LDS kernel,
// factor + inverse), panel solve as a GEMM with the inverse, and the rest of
// the panel updated with the slice.
template <typename T>
hipError_t factor_panel(CULAPACKhandle h, hipStream_t s, bool lower, long n, T *A, long lda, T *W,
                        long j, long jb) {
  // panel slice width: one diagonal block in LDS
  constexpr int IB = hc::diag_max<T>();
  const T one = hc::from_real<T>(1.0), mone = hc::from_real<T>(-1.0), zero = hc::from_real<T>(0.0);
  auto at = [&](long r, long c) { return A + c * lda + r; };
  const char uplo = lower ? 'L' : 'U';
  for (long jj = j; jj < j + jb; jj += IB) {
    const int ib = (int)std::min<long>(IB, j + jb - jj);
    HC_TRY(hc::diag_potrf<T>(s, uplo, ib, at(jj, jj), lda, W, IB, h->dinfo, jj));
    const long r0 = jj + ib, rest = n - r0;
    hc::GemmOpts o;
    o.abort_flag = h->dinfo;
    if (rest > 0) {
      if (lower) {  // L21 <- A21 inv(L11)^H, in place: a tile spans all ib columns
        o.min_nt = ib;
        HC_TRY(hc::gemm<T>(s, hc::OP_N, conj_op<T>(), rest, ib, ib, one, at(r0, jj), lda, W, IB, zero,
                           nullptr, 0, at(r0, jj), lda, o));
      } else {      // U12 <- inv(U11^H) A12, in place: a tile spans all ib rows
        o.min_mt = ib;
        HC_TRY(hc::gemm<T>(s, hc::OP_N, hc::OP_N, ib, rest, ib, one, W, IB, at(jj, r0), lda, zero,
                           nullptr, 0, at(jj, r0), lda, o));
      }
    }
    const long rem = j + jb - r0;  // panel columns still to update
    if (rem > 0) {
      hc::GemmOpts u;
      u.abort_flag = h->dinfo;
      if (lower) {
        u.tri_out = hc::TRI_LOWER;
        HC_TRY(hc::gemm<T>(s, hc::OP_N, conj_op<T>(), rest, rem, ib, mone, at(r0, jj), lda,
                           at(r0, jj), lda, one, at(r0, r0), lda, at(r0, r0), lda, u));
      } else {
        u.tri_out = hc::TRI_UPPER;
        HC_TRY(hc::gemm<T>(s, conj_op<T>(), hc::OP_N, rem, rest, ib, mone, at(jj, r0), lda,
                           at(jj, r0), lda, one, at(r0, r0), lda, at(r0, r0), lda, u));
      }
    }
  }
  return hipSuccess;
}

// Right-looking blocked Cholesky with look-ahead (the reference's loop is
// left-looking, lapack/dpotrf.c:248-286; both produce the same factor).
// With panel j factored, the main stream applies it to the next panel's
// columns only (complex: the look-ahead stream does, see below) and runs the
// large trailing SYRK/HERK on everything beyond; the look-ahead stream
// factors the next panel.  The two streams touch disjoint column ranges, and
// each panel is joined back into the main stream before its own trailing
// update.
// Event pair around one trailing update when the handle's timing is on; the
// event pool grows on demand (outside any timed work it is already sized).
hipError_t timing_begin(CULAPACKhandle h, hipStream_t s, double flops) {
  if (!h->timing) return hipSuccess;
  if (h->tcount == h->tcap) {
    const long cap = h->tcap ? 2 * h->tcap : 256;
    auto *ev = static_cast<hipEvent_t *>(std::realloc(h->tev, 2 * cap * sizeof(hipEvent_t)));
    if (ev == nullptr) return hipErrorOutOfMemory;
    h->tev = ev;
    auto *fl = static_cast<double *>(std::realloc(h->tflops, cap * sizeof(double)));
    if (fl == nullptr) return hipErrorOutOfMemory;
    h->tflops = fl;
    for (long i = 2 * h->tcap; i < 2 * cap; i++) {
      hipError_t e = hipEventCreate(&h->tev[i]);
      if (e != hipSuccess) {
        h->tcap = i / 2;
        return e;
      }
    }
    h->tcap = cap;
  }
  h->tflops[h->tcount] = flops;
  return hipEventRecord(h->tev[2 * h->tcount], s);
}

hipError_t timing_end(CULAPACKhandle h, hipStream_t s) {
  if (!h->timing) return hipSuccess;
  return hipEventRecord(h->tev[2 * h->tcount++ + 1], s);
}

// bf16 (FP32 data, lower only): after each panel is factored it is also
// converted to bfloat16 with k contiguous (two alternating buffers at lp_ws,
// potrf_bf16_ws_bytes) and the next-panel and trailing updates take their
// products from those copies on the BF16 matrix cores.  split: each copy is
// the pair of rows [hi | hi | lo] and [hi | lo | hi] (x = hi + lo), and the
// products run over 3 ldt: three BF16 products, ~16 significant bits.
// Workgroups of the FP32 copy made beside the first panel (four per CU).
constexpr long kConvertWorkgroups = 1024;

// Flags and workspace of the one-launch panel factorisation, allocated once
// per handle (the flags zeroed then; every launch uses a new epoch).
hipError_t panel_resources(CULAPACKhandle h) {
  if (h->pflags != nullptr) return hipSuccess;
  const size_t flag_bytes = (size_t)hc::kPanelFlagInts * sizeof(int);
  const size_t ws_bytes = (size_t)hc::panel_ws_elems(hc::kPanelMaxCols) * sizeof(double);
  void *p = nullptr;
  HC_TRY(hipMalloc(&p, flag_bytes + ws_bytes));
  const hipError_t e = hipMemset(p, 0, flag_bytes);
  if (e != hipSuccess) {
    (void)hipFree(p);
    return e;
  }
  h->pflags = static_cast<int *>(p);
  h->pws = static_cast<char *>(p) + flag_bytes;
  h->pepoch = 0;
  return hipSuccess;
}

template <typename T>
hipError_t potrf_device(CULAPACKhandle h, char uplo, long n, T *A, long lda, T *W, bool bf16 = false,
                        void *lp_ws = nullptr, bool split = false, float f16_scale = 0.0f,
                        const hc::PreConvert *pre = nullptr) {
  const long NB = potrf_nb(h, prec_of<T>(bf16), n);
  hipStream_t s0 = h->stream, s1 = h->side;
  const bool lower = (uplo == 'L');
  const T one = hc::from_real<T>(1.0), mone = hc::from_real<T>(-1.0);
  auto at = [&](long r, long c) { return A + c * lda + r; };
  const long ldt = (NB + 63) / 64 * 64;
  split = split && bf16;
  // FP16 copies of the panels scaled by f16_scale (2^e); the products carry
  // f16_scale^2, undone exactly by alpha = -f16_scale^-2
  const bool f16 = bf16 && !split && f16_scale > 0.0f;
  const float lp_alpha = f16 ? -1.0f / (f16_scale * f16_scale) : -1.0f;
  const long kk = split ? 3 * ldt : ldt;  // K of the bf16 products (and row stride of the copies)
  // per panel buffer: the A-side copy, then (split) the B-side copy
  unsigned short *Pt[2], *Qt[2];
  for (int b = 0; b < 2; b++) {
    Pt[b] = static_cast<unsigned short *>(lp_ws) + (lp_ws ? (split ? 2 : 1) * b * n * kk : 0);
    Qt[b] = split ? Pt[b] + n * kk : Pt[b];
  }
  if (bf16 && (!lower || lp_ws == nullptr || !std::is_same<T, float>::value)) return hipErrorInvalidValue;
  auto convert = [&](hipStream_t s, long m, long kp, const T *P, int b, long col0) -> hipError_t {
    if constexpr (std::is_same<T, float>::value) {
      if (f16) return hc::panel_to_f16(s, m, kp, P, lda, Pt[b], ldt, f16_scale, h->dinfo, col0);
      return split ? hc::panel_to_bf16_split(s, m, kp, P, lda, Pt[b], Qt[b], ldt)
                   : hc::panel_to_bf16(s, m, kp, P, lda, Pt[b], ldt);
    }
    return hipErrorInvalidValue;
  };
  // Lower real panels in one launch (panel_potrf) when the tuning table asks
  // for it; by default a panel is the chain of diagonal-block and
  // panel-product launches (factor_panel).  Measured on MI355X the one-launch
  // panel is slower beside the trailing update: its resident workgroups keep
  // the CUs the 16-bit update's workgroups need (one per CU, all of the LDS
  // and registers), and its left-looking steps put two 128^3 products on one
  // CU on the critical path (dpotrf n = 16384 34.5 ms against 33.5 ms; FP16
  // solve n = 32768 59.8 ms against 57.6 ms).
  const bool one_launch = lower && !hc::is_complex<T>::value && NB <= hc::kPanelMaxCols &&
                          hc::tuned_panel_launch(prec_of<T>(bf16), false);
  if (one_launch) HC_TRY(panel_resources(h));
  auto panel = [&](hipStream_t s, long jj, long jbb) -> hipError_t {
    if constexpr (!hc::is_complex<T>::value) {
      if (one_launch)
        return hc::panel_potrf<T>(s, n - jj, jbb, at(jj, jj), lda, static_cast<T *>(h->pws), h->pflags,
                                  ++h->pepoch, h->dinfo, jj);
    }
    return factor_panel<T>(h, s, lower, n, A, lda, W, jj, jbb);
  };
  if (pre != nullptr) {
    // FP32 copy of A made here: the first block column before the first
    // panel, the rest (and the norm) on the look-ahead stream beside it; the
    // main stream waits for it before the first update beyond that panel.
    if constexpr (!std::is_same<T, float>::value) return hipErrorInvalidValue;
    else {
      const long csplit = std::min(n, (std::min(NB, n) + 63) / 64 * 64);
      HC_TRY(hc::to_float_norm_cols(s0, pre->uplo, n, pre->A, pre->lda, A, lda, pre->overflow, pre->ws, 0,
                                    csplit, 0));
      HC_TRY(hipEventRecord(h->ev_update, s0));
      HC_TRY(hipStreamWaitEvent(s1, h->ev_update, 0));
      // a bounded grid, so that the panel's launches find room beside it
      HC_TRY(hc::to_float_norm_cols(s1, pre->uplo, n, pre->A, pre->lda, A, lda, pre->overflow, pre->ws, csplit,
                                    n, kConvertWorkgroups));
      HC_TRY(hc::to_float_norm_finish(s1, n, pre->ws, pre->anorm));
      HC_TRY(hipEventRecord(h->ev_panel, s1));
    }
  }
  HC_TRY(panel(s0, 0, std::min(NB, n)));
  if (bf16) HC_TRY(convert(s0, n, std::min(NB, n), A, 0, 0));
  if (pre != nullptr) HC_TRY(hipStreamWaitEvent(s0, h->ev_panel, 0));
  for (long j = 0; j < n; j += NB) {
    const int pj = (int)((j / NB) & 1);
    const long jb = std::min(NB, n - j);
    const long jn = j + jb;                 // next panel
    if (jn >= n) break;
    const long jbn = std::min(NB, n - jn);
    hc::GemmOpts u;
    u.abort_flag = h->dinfo;
    u.bf16 = bf16;
    // The next panel's update needs this panel and the previous trailing
    // update; the trailing update beyond the next panel needs only those
    // too.  On the look-ahead stream (ahead of the next panel's
    // factorisation) it runs beside the trailing update instead of before
    // it.  Measured on MI355X that pays for the complex factorisation
    // (zpotrf n = 8192: 20.1 -> 19.9 ms) but not for the real one, whose
    // look-ahead chain is the longer branch late in the factorisation and
    // then starts behind a slowed update (dpotrf n = 16384: 33.5 -> 34.2 ms;
    // 33.6-33.8 ms when only the early, long trailing updates get it); the
    // FP32 / 16-bit factorisation's look-ahead stream is already the busier.
    const bool upd_side = !bf16 && hc::is_complex<T>::value;
    hipStream_t su = s0;
    if (upd_side) {
      HC_TRY(hipEventRecord(h->ev_update, s0));
      HC_TRY(hipStreamWaitEvent(s1, h->ev_update, 0));
      su = s1;
    }
    if (lower) {  // next panel columns: A[jn:, jn:jn+jbn] -= P[jn:] P[jn:jn+jbn]^H
      u.tri_out = hc::TRI_LOWER;
      bool done = false;
      if constexpr (std::is_same<T, float>::value) {
        if (bf16) {
          const unsigned short *Pn = Pt[pj] + (jn - j) * kk, *Qn = Qt[pj] + (jn - j) * kk;
          HC_TRY(hc::gemm_bf16_nt(su, n - jn, jbn, kk, lp_alpha, Pn, kk, Qn, kk, 1.0f, at(jn, jn), lda,
                                  at(jn, jn), lda, hc::TRI_LOWER, false, h->dinfo, f16));
          done = true;
        }
      }
      if (!done)
        HC_TRY(hc::gemm<T>(su, hc::OP_N, conj_op<T>(), n - jn, jbn, jb, mone, at(jn, j), lda, at(jn, j),
                           lda, one, at(jn, jn), lda, at(jn, jn), lda, u));
    } else {
      u.tri_out = hc::TRI_UPPER;
      HC_TRY(hc::gemm<T>(su, conj_op<T>(), hc::OP_N, jbn, n - jn, jb, mone, at(j, jn), lda, at(j, jn),
                         lda, one, at(jn, jn), lda, at(jn, jn), lda, u));
    }
    if (!upd_side) {
      HC_TRY(hipEventRecord(h->ev_update, s0));
      HC_TRY(hipStreamWaitEvent(s1, h->ev_update, 0));
    }
    HC_TRY(panel(s1, jn, jbn));
    if (bf16) HC_TRY(convert(s1, n - jn, jbn, at(jn, jn), pj ^ 1, jn));
    HC_TRY(hipEventRecord(h->ev_panel, s1));
    const long t0 = jn + jbn, t = n - t0;   // trailing matrix beyond the next panel
    if (t > 0) {
      const double flops = (hc::is_complex<T>::value ? 4.0 : 1.0) * (double)t * (double)(t + 1) * (double)jb;
      HC_TRY(timing_begin(h, s0, flops));
      bool done = false;
      if constexpr (std::is_same<T, float>::value) {
        if (bf16) {
          const unsigned short *Pn = Pt[pj] + (t0 - j) * kk, *Qn = Qt[pj] + (t0 - j) * kk;
          HC_TRY(hc::gemm_bf16_nt(s0, t, t, kk, lp_alpha, Pn, kk, Qn, kk, 1.0f, at(t0, t0), lda,
                                  at(t0, t0), lda, hc::TRI_LOWER, true, h->dinfo, f16));
          done = true;
        }
      }
      // late, shorter trailing updates leave room for the look-ahead chain
      const int max_wgs = hc::tuned_trailing_cap(prec_of<T>(bf16), t);
      if (done) {
      } else if (lower)
        HC_TRY(hc::herk<T>(s0, 'L', 'N', t, jb, -1.0, at(t0, j), lda, 1.0, at(t0, t0), lda, h->dinfo,
                           bf16, max_wgs));
      else
        HC_TRY(hc::herk<T>(s0, 'U', 'C', t, jb, -1.0, at(j, t0), lda, 1.0, at(t0, t0), lda, h->dinfo,
                           bf16, max_wgs));
      HC_TRY(timing_end(h, s0));
    }
    HC_TRY(hipStreamWaitEvent(s0, h->ev_panel, 0));
  }
  return hipSuccess;
}

// Recursive triangular inverse: inv([A11 0; A21 A22]) =
// [X11 0; -X22 A21 X11  X22] (lower; upper is the mirror).  The reference's
// loop (lapack/dtrtri.c:360-458) walks nb-column blocks; halving instead keeps
// the two TRMMs per level large.  X holds an n2 x n1 intermediate.
// The leaves (IB x IB diagonal blocks at multiples of IB: every split point
// is one) only read their own block, which no product of the recursion
// writes, so they are all inverted first, in one launch.
template <typename T>
hipError_t trtri_rec(hipStream_t s, char uplo, char diag, long n, T *A, long lda, T *X, bool top = true) {
  constexpr int IB = hc::diag_max<T>();
  if (n <= IB) return top ? hc::diag_trtri<T>(s, uplo, diag, (int)n, A, lda, A, lda, nullptr, 0, nullptr)
                          : hipSuccess;
  if (top) HC_TRY(hc::diag_trtri_blocks<T>(s, uplo, diag, n, IB, A, lda));
  const long n1 = ((n / 2 + IB - 1) / IB) * IB, n2 = n - n1;
  T *A11 = A, *A22 = A + n1 * lda + n1;
  HC_TRY(trtri_rec(s, uplo, diag, n1, A11, lda, X, false));
  HC_TRY(trtri_rec(s, uplo, diag, n2, A22, lda, X, false));
  const T one = hc::from_real<T>(1.0), mone = hc::from_real<T>(-1.0);
  if (uplo == 'L') {
    T *A21 = A + n1;
    HC_TRY(hc::trmm<T>(s, 'L', 'L', 'N', diag, n2, n1, mone, A22, lda, A21, lda, X, n2));
    return hc::trmm<T>(s, 'R', 'L', 'N', diag, n2, n1, one, A11, lda, X, n2, A21, lda);
  }
  T *A12 = A + n1 * lda;
  HC_TRY(hc::trmm<T>(s, 'L', 'U', 'N', diag, n1, n2, mone, A11, lda, A12, lda, X, n1));
  return hc::trmm<T>(s, 'R', 'U', 'N', diag, n1, n2, one, A22, lda, X, n1, A12, lda);
}

// Recursive triangular square: lower L^H L =
// [L11^H L11 + L21^H L21, .; L22^H L21, L22^H L22]; upper U U^H mirrors it.
// The reference's loop is lapack/dlauum.c:188-281.
// A leaf (IB x IB diagonal block at a multiple of IB) still holds its factor
// block when the recursion reaches it, but the products above it read that
// block until then, so the leaves are squared beforehand into `leaves`
// (block b at leaves + b IB^2, one launch) and only copied into place here.
template <typename T>
hipError_t lauum_rec(hipStream_t s, char uplo, long n, T *A, long lda, T *X, const T *leaves) {
  constexpr int IB = hc::diag_max<T>();
  if (n <= IB) return hc::tri_copy<T>(s, uplo, (int)n, leaves, IB, A, lda);
  const long n1 = ((n / 2 + IB - 1) / IB) * IB, n2 = n - n1;
  T *A11 = A, *A22 = A + n1 * lda + n1;
  const T one = hc::from_real<T>(1.0);
  HC_TRY(lauum_rec(s, uplo, n1, A11, lda, X, leaves));
  if (uplo == 'L') {
    T *A21 = A + n1;
    HC_TRY(hc::herk<T>(s, 'L', 'C', n1, n2, 1.0, A21, lda, 1.0, A11, lda));
    HC_TRY(hc::trmm<T>(s, 'L', 'L', 'C', 'N', n2, n1, one, A22, lda, A21, lda, X, n2));
    HC_TRY(hc::copy2d<T>(s, n2, n1, X, n2, A21, lda));
  } else {
    T *A12 = A + n1 * lda;
    HC_TRY(hc::herk<T>(s, 'U', 'N', n1, n2, 1.0, A12, lda, 1.0, A11, lda));
    HC_TRY(hc::trmm<T>(s, 'R', 'U', 'C', 'N', n1, n2, one, A22, lda, A12, lda, X, n1));
    HC_TRY(hc::copy2d<T>(s, n1, n2, X, n1, A12, lda));
  }
  return lauum_rec(s, uplo, n2, A22, lda, X, leaves + (n1 / IB) * (long)IB * IB);
}

template <typename T>
size_t rec_ws_bytes(long n) {
  constexpr int IB = hc::diag_max<T>();
  if (n <= IB) return 0;
  const long n1 = ((n / 2 + IB - 1) / IB) * IB;
  return (size_t)n1 * (size_t)(n - n1) * sizeof(T);
}

// Copies the device info word back and waits for the handle's streams.
CUresult finish(CULAPACKhandle h, long *info) {
  HC_CHECK(hipMemcpyAsync(h->hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HC_CHECK(hipStreamSynchronize(h->stream));
  *info = (long)h->hinfo[0];
  if (h->hinfo[0] == hc::kInfoAbandoned) {  // a dependency wait of the panel kernel gave up
    *info = 0;
    return CUDA_ERROR_LAUNCH_FAILED;
  }
  return CUDA_SUCCESS;
}

template <typename T>
CUresult potrf_api(const char *name, CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A,
                   size_t lda, long *info) {
  // reference lapack/dpotrf.c:173-182
  if (info == nullptr) return bad_arg(name, 6);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -2; return bad_arg(name, 2); }
  if (lda < n) { *info = -4; return bad_arg(name, 4); }
  if (n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  void *ws = nullptr;
  HC_CHECK(workspace(h, (size_t)hc::diag_max<T>() * hc::diag_max<T>() * sizeof(T), &ws));
  HC_CHECK(hipMemsetAsync(h->dinfo, 0, sizeof(int), h->stream));
  HC_CHECK(potrf_device<T>(h, (char)uplo, (long)n, dptr<T>(A), (long)lda, static_cast<T *>(ws)));
  return finish(h, info);
}

template <typename T>
CUresult trtri_api(const char *name, CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n,
                   CUdeviceptr A, size_t lda, long *info) {
  // reference lapack/dtrtri.c:331-340 (info = -5 for lda)
  if (info == nullptr) return bad_arg(name, 7);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -2; return bad_arg(name, 2); }
  if (!hc::valid_diag(diag)) { *info = -3; return bad_arg(name, 3); }
  if (lda < n) { *info = -5; return bad_arg(name, 5); }
  if (n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  T *Ap = dptr<T>(A);
  if (diag == CBlasNonUnit) {
    HC_CHECK(hc::diag_zero_scan<T>(h->stream, (char)uplo, (long)n, Ap, (long)lda, h->dinfo));
    CUresult r = finish(h, info);
    if (r != CUDA_SUCCESS || *info != 0) return r;
  }
  void *ws = nullptr;
  HC_CHECK(workspace(h, rec_ws_bytes<T>((long)n) + 256, &ws));
  HC_CHECK(trtri_rec<T>(h->stream, (char)uplo, (char)diag, (long)n, Ap, (long)lda, static_cast<T *>(ws)));
  HC_CHECK(hipStreamSynchronize(h->stream));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult lauum_api(const char *name, CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A,
                   size_t lda, long *info) {
  // reference lapack/dlauum.c:159-168
  if (info == nullptr) return bad_arg(name, 6);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -2; return bad_arg(name, 2); }
  if (lda < n) { *info = -4; return bad_arg(name, 4); }
  if (n == 0) return CUDA_SUCCESS;
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  void *ws = nullptr;
  constexpr int IB = hc::diag_max<T>();
  if ((long)n <= IB) {
    HC_CHECK(hc::diag_lauum<T>(h->stream, (char)uplo, (int)n, dptr<T>(A), (long)lda));
  } else {
    const size_t xb = (rec_ws_bytes<T>((long)n) + 255) / 256 * 256;
    const size_t lb = (size_t)((n + IB - 1) / IB) * IB * IB * sizeof(T);
    HC_CHECK(workspace(h, xb + lb, &ws));
    T *leaves = reinterpret_cast<T *>(static_cast<char *>(ws) + xb);
    HC_CHECK(hc::diag_lauum_blocks<T>(h->stream, (char)uplo, (long)n, IB, dptr<const T>(A), (long)lda, leaves));
    HC_CHECK(lauum_rec<T>(h->stream, (char)uplo, (long)n, dptr<T>(A), (long)lda, static_cast<T *>(ws), leaves));
  }
  HC_CHECK(hipStreamSynchronize(h->stream));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult potri_api(const char *name, CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A,
                   size_t lda, long *info) {
  // reference lapack/dpotri.c:25-46: TRTRI (non-unit) then LAUUM
  if (info == nullptr) return bad_arg(name, 6);
  *info = 0;
  if (h == nullptr) { *info = -1; return bad_arg(name, 1); }
  if (!hc::valid_uplo(uplo)) { *info = -2; return bad_arg(name, 2); }
  if (lda < n) { *info = -4; return bad_arg(name, 4); }
  if (n == 0) return CUDA_SUCCESS;
  CUresult r = trtri_api<T>(name, h, uplo, CBlasNonUnit, n, A, lda, info);
  if (r != CUDA_SUCCESS || *info != 0) return r;
  return lauum_api<T>(name, h, uplo, n, A, lda, info);
}

template <typename T>
CUresult trti22_api(const char *name, CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n,
                    CUdeviceptr A, size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr info,
                    CUstream stream) {
  // reference lapack/dtrtri.c:301-324 (there n <= 32; here one LDS block)
  if (h == nullptr) return bad_arg(name, 1);
  if (!hc::valid_uplo(uplo)) return bad_arg(name, 2);
  if (!hc::valid_diag(diag)) return bad_arg(name, 3);
  if (n > (size_t)hc::diag_max<T>()) return CUDA_ERROR_INVALID_VALUE;
  if (lda < n) return bad_arg(name, 6);
  if (ldb < n) return bad_arg(name, 8);
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  int *dinfo = dptr<int>(info);
  if (n == 0) {
    if (dinfo != nullptr) HC_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int), stream));
    return CUDA_SUCCESS;
  }
  if (dinfo != nullptr) {
    if (diag == CBlasNonUnit)
      HC_CHECK(hc::diag_zero_scan<T>(stream, (char)uplo, (long)n, dptr<const T>(A), (long)lda, dinfo));
    else
      HC_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int), stream));
  }
  HC_CHECK(hc::diag_trtri<T>(stream, (char)uplo, (char)diag, (int)n, dptr<const T>(A), (long)lda,
                             dptr<T>(B), (long)ldb, nullptr, 0, dinfo));
  return CUDA_SUCCESS;
}

template <typename T>
CUresult logdet_api(const char *name, CULAPACKhandle h, CUdeviceptr x, size_t incx, size_t n,
                    double *result, CUstream stream) {
  // reference lapack/dlogdet.c:37-74
  if (result == nullptr) return bad_arg(name, 5);
  if (n == 0) { *result = 0.0; return CUDA_SUCCESS; }
  if (h == nullptr) return bad_arg(name, 1);
  if (incx == 0) return bad_arg(name, 3);
  hc::DeviceGuard g(h->device);
  HC_CHECK(g.status());
  // Partial sums and the result sit behind the info word in the handle's
  // small device buffer area: a separate allocation would be per call.
  void *ws = nullptr;
  HC_CHECK(workspace(h, (2 * hc::kLogdetMaxBlocks + 8) * sizeof(double), &ws));
  double *partials = static_cast<double *>(ws);
  double *out = partials + 2 * hc::kLogdetMaxBlocks;
  // The handle workspace is ordered on the handle stream; join the caller's.
  HC_CHECK(hipStreamSynchronize(h->stream));
  HC_CHECK(hc::logdet<T>(stream, dptr<const T>(x), (long)incx, (long)n, partials, out));
  double *hres = reinterpret_cast<double *>(h->hinfo) + 1;
  HC_CHECK(hipMemcpyAsync(hres, out, sizeof(double), hipMemcpyDeviceToHost, stream));
  HC_CHECK(hipStreamSynchronize(stream));
  *result = *hres;
  return CUDA_SUCCESS;
}

}  // namespace

namespace hc {

hipError_t lapack_workspace(CULAPACKhandle h, size_t bytes, void **out) {
  return workspace(h, bytes, out);
}

template <typename T>
hipError_t potrf_on_device(CULAPACKhandle h, char uplo, long n, T *A, long lda, T *W, bool bf16,
                           void *lp_ws, bool split, float f16_scale, const PreConvert *pre) {
  return potrf_device<T>(h, uplo, n, A, lda, W, bf16, lp_ws, split, f16_scale, pre);
}
template hipError_t potrf_on_device<double>(CULAPACKhandle, char, long, double *, long, double *, bool,
                                            void *, bool, float, const PreConvert *);
template hipError_t potrf_on_device<float>(CULAPACKhandle, char, long, float *, long, float *, bool,
                                           void *, bool, float, const PreConvert *);

size_t potrf_bf16_ws_bytes(CULAPACKhandle h, long n, bool split) {
  const long ldt = (potrf_nb(h, 'B', n) + 63) / 64 * 64;
  // two alternating panel buffers of n x ldt bf16; split: A- and B-side copies of n x 3 ldt each
  return 2 * (size_t)n * (size_t)ldt * 2 * (split ? 6 : 1);
}

}  // namespace hc

extern "C" {

CUresult cuDpotrf(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potrf_api<double>("cuDpotrf", h, uplo, n, A, lda, info);
}
CUresult cuLAPACKSetBlockSize(CULAPACKhandle h, char precision, long nb) {
  if (h == nullptr) return bad_arg("cuLAPACKSetBlockSize", 1);
  if (precision != 'D' && precision != 'Z' && precision != 'S' && precision != 'B')
    return bad_arg("cuLAPACKSetBlockSize", 2);
  if (nb < 0 || (nb > 0 && nb < hc::diag_max<double>()) || nb > 8192)
    return bad_arg("cuLAPACKSetBlockSize", 3);
  h->nb_override[prec_slot(precision)] = nb;
  return CUDA_SUCCESS;
}

CUresult cuSpotrf(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potrf_api<float>("cuSpotrf", h, uplo, n, A, lda, info);
}
CUresult cuZpotrf(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potrf_api<zd>("cuZpotrf", h, uplo, n, A, lda, info);
}
CUresult cuDtrtri(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A,
                  size_t lda, long *info) {
  return trtri_api<double>("cuDtrtri", h, uplo, diag, n, A, lda, info);
}
CUresult cuZtrtri(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A,
                  size_t lda, long *info) {
  return trtri_api<zd>("cuZtrtri", h, uplo, diag, n, A, lda, info);
}
CUresult cuDtrti22(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A,
                   size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr info, CUstream stream) {
  return trti22_api<double>("cuDtrti22", h, uplo, diag, n, A, lda, B, ldb, info, stream);
}
CUresult cuZtrti22(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A,
                   size_t lda, CUdeviceptr B, size_t ldb, CUdeviceptr info, CUstream stream) {
  return trti22_api<zd>("cuZtrti22", h, uplo, diag, n, A, lda, B, ldb, info, stream);
}
CUresult cuDlauum(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return lauum_api<double>("cuDlauum", h, uplo, n, A, lda, info);
}
CUresult cuZlauum(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return lauum_api<zd>("cuZlauum", h, uplo, n, A, lda, info);
}
CUresult cuDpotri(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potri_api<double>("cuDpotri", h, uplo, n, A, lda, info);
}
CUresult cuZpotri(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potri_api<zd>("cuZpotri", h, uplo, n, A, lda, info);
}
CUresult cuDlogdet(CULAPACKhandle h, CUdeviceptr x, size_t incx, size_t n, double *result,
                   CUstream stream) {
  return logdet_api<double>("cuDlogdet", h, x, incx, n, result, stream);
}
CUresult cuZlogdet(CULAPACKhandle h, CUdeviceptr x, size_t incx, size_t n, double *result,
                   CUstream stream) {
  return logdet_api<zd>("cuZlogdet", h, x, incx, n, result, stream);
}

// Single precision and single complex: level-3 products on the FP32 matrix
// cores, diagonal blocks computed in double / double complex (reference
// lapack/cpotrf.c:174, lapack/strtri.c:301,369, lapack/ctrtri.c:301,326,
// lapack/slauum.c:197, lapack/clauum.c:154, lapack/spotri.c:25,
// lapack/cpotri.c:25, lapack/slogdet.c:37, lapack/clogdet.c:38).
CUresult cuCpotrf(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potrf_api<hc::zf>("cuCpotrf", h, uplo, n, A, lda, info);
}
CUresult cuStrtri(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A, size_t lda,
                  long *info) {
  return trtri_api<float>("cuStrtri", h, uplo, diag, n, A, lda, info);
}
CUresult cuCtrtri(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A, size_t lda,
                  long *info) {
  return trtri_api<hc::zf>("cuCtrtri", h, uplo, diag, n, A, lda, info);
}
CUresult cuSlauum(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return lauum_api<float>("cuSlauum", h, uplo, n, A, lda, info);
}
CUresult cuClauum(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return lauum_api<hc::zf>("cuClauum", h, uplo, n, A, lda, info);
}
CUresult cuSpotri(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potri_api<float>("cuSpotri", h, uplo, n, A, lda, info);
}
CUresult cuCpotri(CULAPACKhandle h, CBlasUplo uplo, size_t n, CUdeviceptr A, size_t lda, long *info) {
  return potri_api<hc::zf>("cuCpotri", h, uplo, n, A, lda, info);
}
CUresult cuStrti22(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A, size_t lda,
                   CUdeviceptr B, size_t ldb, CUdeviceptr info, CUstream stream) {
  return trti22_api<float>("cuStrti22", h, uplo, diag, n, A, lda, B, ldb, info, stream);
}
CUresult cuCtrti22(CULAPACKhandle h, CBlasUplo uplo, CBlasDiag diag, size_t n, CUdeviceptr A, size_t lda,
                   CUdeviceptr B, size_t ldb, CUdeviceptr info, CUstream stream) {
  return trti22_api<hc::zf>("cuCtrti22", h, uplo, diag, n, A, lda, B, ldb, info, stream);
}
CUresult cuSlogdet(CULAPACKhandle h, CUdeviceptr x, size_t incx, size_t n, float *result, CUstream stream) {
  if (result == nullptr) return bad_arg("cuSlogdet", 5);
  double r = 0.0;
  const CUresult st = logdet_api<float>("cuSlogdet", h, x, incx, n, &r, stream);
  if (st == CUDA_SUCCESS) *result = (float)r;
  return st;
}
CUresult cuClogdet(CULAPACKhandle h, CUdeviceptr x, size_t incx, size_t n, float *result, CUstream stream) {
  if (result == nullptr) return bad_arg("cuClogdet", 5);
  double r = 0.0;
  const CUresult st = logdet_api<hc::zf>("cuClogdet", h, x, incx, n, &r, stream);
  if (st == CUDA_SUCCESS) *result = (float)r;
  return st;
}

}  // extern "C"
