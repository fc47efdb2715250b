// gemm_mixed.hpp — split-FP16 MFMA subspace products (see gemm_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace kukan {
// A band block X (R x M, column-major) scaled by s and split once into the fragment layouts
// of both products: as the first operand of X^T B (TN) and as the left factor of X C (NN).
// The buffers are the ones the products split into themselves, so the split is valid until
// the next product that splits its own operand.
struct SplitBands {
  const double *X = nullptr;
  long ldx = 0, R = 0;
  int M = 0;
  const double *s = nullptr;
  _Float16 *th = nullptr, *tl = nullptr, *nh = nullptr, *nl = nullptr;
};
void split_bands(const double *X, long ldx, long R, int M, const double *s, SplitBands &out,
                 hipStream_t st);
// scale[j] = 2^(14 - e_j) with max_i |A(i,j)| < 2^e_j (1 for a zero column)
void colscale(const double *A, long lda, long rows, int cols, double *scale, hipStream_t s);
// the same scales from column maxima given as FP64 bit patterns
void maxscale(const unsigned long long *cmax, int cols, double *scale, hipStream_t s);
// scale[j] = 2^(13 - e_j) with sqrt(g_j / dv) < 2^e_j: a bound for columns with
// dv sum_i A(i,j)^2 = g_j, no pass over A
void normscale(const double *g, int cols, double dv, double *scale, hipStream_t s);
// C(M1 x M2) = alpha A^T B + beta C with column scales sa (A) and sb (B).  tri = 1 / 2 (M1 = M2):
// only the 256 x 256 tiles on and below / above the diagonal are computed and written (for a
// consumer that reads the lower / upper triangle alone)
void gemm_mixed_tn(int M1, int M2, long K, double alpha, const double *A, long lda,
                   const double *sa, const double *B, long ldb, const double *sb, double beta,
                   double *C, long ldc, int nterms, hipStream_t s, const SplitBands *pa = nullptr,
                   int tri = 0);
// Y(R x M2) = alpha X C + beta Y with column scales sx of X (length K).  X is read only by
// the pre-split pass that precedes the product, so Y may be X itself (same ld, beta = 0).
// colsq (beta = 0 only): dv sum_i Y(i,j)^2 of the written Y per column, from the epilogue.
// c_upper: C is upper triangular (zero below the diagonal), so column j needs only k <= j.
void gemm_mixed_nn(long R, int K, int M2, double alpha, const double *X, long ldx,
                   const double *sx, const double *C, long ldc, double beta, double *Y, long ldy,
                   int nterms, hipStream_t s, const SplitBands *px = nullptr,
                   double *colsq = nullptr, double dv = 1.0, bool c_upper = false);
}  // namespace kukan
