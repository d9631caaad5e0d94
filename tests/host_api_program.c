/*
 * A caller of the reference's host-memory API, compiled and linked against
 * libcucholesky.so by tests/test_gpu_host_api.py (it includes only
 * include/lapack.h, as a reference program does).
 *   host_api_program <in> <out>
 * in:  int64 n, then A (n x n SPD), B (n x n), Z (n x n Hermitian PD,
 *      interleaved complex), column-major doubles
 * out: info(dpotrf), L, logdet, info(dpotri), inv(A), info(zpotrf), U (complex),
 *      C = A B^T (dgemm), X = inv(L) B (dtrsm)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lapack.h"

static int rd(void *p, size_t sz, size_t n, FILE *f) { return fread(p, sz, n, f) == n; }

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  FILE *fi = fopen(argv[1], "rb"), *fo = fopen(argv[2], "wb");
  if (!fi || !fo) return 3;
  long long n64 = 0;
  if (!rd(&n64, sizeof n64, 1, fi)) return 4;
  const size_t n = (size_t)n64, nn = n * n;
  double *A = malloc(nn * sizeof(double)), *B = malloc(nn * sizeof(double));
  double complex *Z = malloc(nn * sizeof(double complex));
  if (!rd(A, sizeof(double), nn, fi) || !rd(B, sizeof(double), nn, fi) || !rd(Z, sizeof(double complex), nn, fi))
    return 5;
  long info = -99;
  double *F = malloc(nn * sizeof(double));
  memcpy(F, A, nn * sizeof(double));
  dpotrf(CBlasLower, n, F, n, &info);
  fwrite(&info, sizeof info, 1, fo);
  fwrite(F, sizeof(double), nn, fo);
  const double ld = dlogdet(F, n + 1, n);
  fwrite(&ld, sizeof ld, 1, fo);
  double *Inv = malloc(nn * sizeof(double));
  memcpy(Inv, F, nn * sizeof(double));
  dpotri(CBlasLower, n, Inv, n, &info);
  fwrite(&info, sizeof info, 1, fo);
  fwrite(Inv, sizeof(double), nn, fo);
  zpotrf(CBlasUpper, n, Z, n, &info);
  fwrite(&info, sizeof info, 1, fo);
  fwrite(Z, sizeof(double complex), nn, fo);
  double *C = calloc(nn, sizeof(double));
  dgemm(CBlasNoTrans, CBlasTrans, n, n, n, 1.0, A, n, B, n, 0.0, C, n);
  fwrite(C, sizeof(double), nn, fo);
  dtrsm(CBlasLeft, CBlasLower, CBlasNoTrans, CBlasNonUnit, n, n, 1.0, F, n, B, n);
  fwrite(B, sizeof(double), nn, fo);
  fclose(fo);
  fclose(fi);
  free(A); free(B); free(Z); free(F); free(Inv); free(C);
  return 0;
}
