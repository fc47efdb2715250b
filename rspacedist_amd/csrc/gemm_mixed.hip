// gemm_mixed.hip — split-FP16 MFMA products for the subspace work (gfx950).
//
// An FP64 operand x (scaled per column by a power of two s so that max|x s| ~ 2^14) is
// split into hi = fp16(x s) and lo = fp16(x s - hi), 22 significant bits in all.  A
// product is assembled from NT MFMA products (hi*hi + hi*lo + lo*hi [+ lo*lo]), each
// exact in FP32 before accumulation; dropped terms are ~2^-22 relative.  The products use
// v_mfma_f32_32x32x16_f16 (A lane l: row l%32, k 8*(l/32)..+7; B lane l: col l%32, same k;
// D reg r: row 8*(r/4)+4*(l/32)+r%4, col l%32).
//
//   TN:  C(M1 x M2) = alpha A^T B + beta C, A (K x M1), B (K x M2) column-major, K large.
//        Split-K over work-groups; inside a group FP32 accumulation runs over 256-deep
//        K-chunks and is flushed into FP64 registers; groups write FP64 partials that a
//        second kernel reduces in a fixed order (deterministic).
//   NN:  Y(R x M2) = alpha X C + beta Y, X (R x K) column-major (K = number of bands),
//        C small.  X is loaded straight from HBM in the MFMA operand layout (lanes 0-31
//        read 32 consecutive rows of one column), C is pre-split once.
#include "gemm_mixed.hpp"
#include "kukan_internal.hpp"

#include <algorithm>
#include <cmath>

namespace kukan {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr double TARGET = 16384.0;  // 2^14

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

// f64 -> f16 through FP32 with hardware conversions.  The empty asm keeps the compiler from
// merging the two roundings into one correctly rounded f64 -> f16 conversion, which is a
// long software sequence.  Rounding twice is harmless here: x - hi is exact in FP64
// whatever hi is, so only lo's rounding limits the accuracy of the split.
__device__ inline _Float16 to_half(double x) {
  float f = (float)x;
  asm volatile("" : "+v"(f));
  return (_Float16)f;
}

__device__ inline void split8(const double *v, double s, half8 &hi, half8 &lo) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double x = v[j] * s;
    const _Float16 h = to_half(x);
    hi[j] = h;
    lo[j] = to_half(x - (double)h);
  }
}

// per-column power-of-two scale: s_j = 2^(14 - ceil(log2 max|col_j|)) (1 for a zero column)
__global__ __launch_bounds__(256) void k_colscale(const double *A, long lda, long rows, int cols,
                                                  const double *pre, double *scale) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  if (j >= cols) return;
  const double *c = A + (long)j * lda;
  const double f = pre ? fabs(pre[j]) : 1.0;
  double m = 0.0;
  for (long i = threadIdx.x; i < rows; i += 256) m = fmax(m, fabs(c[i]));
  red[threadIdx.x] = m * f;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mx = red[0];
    double sc = 1.0;
    if (mx > 0.0) {
      int e;
      frexp(mx, &e);  // mx < 2^e
      sc = ldexp(1.0, 14 - e);
    }
    scale[j] = sc;
  }
}

// ----------------------------------------------------------------------------- TN
constexpr int TN_T = 128;    // tile edge
constexpr int TN_KS = 32;    // K per LDS stage
constexpr int TN_PITCH = 40; // halves per LDS row (32 + 8 pad)
constexpr int TN_FLUSH = 8;  // stages per FP32 chunk (256 K)

struct TNArgs {
  int M1, M2;
  long K;
  const double *A, *B;
  long lda, ldb;
  const double *sa, *sb;
  double *part;  // [split][M1*M2] column-major M1 x M2
  long kper;     // K per split (multiple of TN_KS)
  int nsplit, tiles_i, tiles_j, nsplit_pad;
};

template <int NT>
__global__ __launch_bounds__(256) void k_gemm_tn(TNArgs p) {
  __shared__ _Float16 sAh[TN_T * TN_PITCH], sAl[TN_T * TN_PITCH];
  __shared__ _Float16 sBh[TN_T * TN_PITCH], sBl[TN_T * TN_PITCH];
  const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
  const int tiles = p.tiles_i * p.tiles_j;
  const int split = (jj / tiles) * 8 + xcd;
  if (split >= p.nsplit) return;
  const int tile = jj % tiles;
  const int ti = tile % p.tiles_i, tj = tile / p.tiles_i;
  const int i0 = ti * TN_T, j0 = tj * TN_T;
  const long kbeg = (long)split * p.kper, kend = min(p.K, kbeg + p.kper);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = w & 1, wj = w >> 1;
  f16v acc[2][2];
  double acc64[2][2][16];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        acc[a][b][r] = 0.f;
        acc64[a][b][r] = 0.0;
      }
    }
  int stage = 0;
  for (long k0 = kbeg; k0 < kend; k0 += TN_KS, stage++) {
    // global -> LDS: 512 units (128 columns x 4 k-groups of 8) per matrix
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int unit = t + 256 * u;
      const int col = unit >> 2, kg = unit & 3;
      const long kk = k0 + 8 * kg;
      double va[8], vb[8];
      const int ia = i0 + col, jb = j0 + col;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const bool kin = kk + q < kend;
        va[q] = (kin && ia < p.M1) ? p.A[(long)ia * p.lda + kk + q] : 0.0;
        vb[q] = (kin && jb < p.M2) ? p.B[(long)jb * p.ldb + kk + q] : 0.0;
      }
      half8 h, l;
      split8(va, ia < p.M1 ? p.sa[ia] : 1.0, h, l);
      *(half8 *)&sAh[col * TN_PITCH + 8 * kg] = h;
      *(half8 *)&sAl[col * TN_PITCH + 8 * kg] = l;
      split8(vb, jb < p.M2 ? p.sb[jb] : 1.0, h, l);
      *(half8 *)&sBh[col * TN_PITCH + 8 * kg] = h;
      *(half8 *)&sBl[col * TN_PITCH + 8 * kg] = l;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; s++) {
      half8 ah[2], al[2], bh[2], bl[2];
      const int ko = 16 * s + 8 * (lane >> 5);
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int row = 64 * wi + 32 * a + (lane & 31);
        ah[a] = *(const half8 *)&sAh[row * TN_PITCH + ko];
        al[a] = *(const half8 *)&sAl[row * TN_PITCH + ko];
      }
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int col = 64 * wj + 32 * b + (lane & 31);
        bh[b] = *(const half8 *)&sBh[col * TN_PITCH + ko];
        bl[b] = *(const half8 *)&sBl[col * TN_PITCH + ko];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
          if (NT == 4)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bl[b], acc[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
    if ((stage + 1) % TN_FLUSH == 0) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int r = 0; r < 16; r++) {
            acc64[a][b][r] += (double)acc[a][b][r];
            acc[a][b][r] = 0.f;
          }
    }
  }
  double *out = p.part + (long)split * p.M1 * p.M2;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = i0 + 64 * wi + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = j0 + 64 * wj + 32 * b + (lane & 31);
        if (m < p.M1 && n < p.M2) out[m + (long)p.M1 * n] = acc64[a][b][r] + (double)acc[a][b][r];
      }
}

__global__ void k_tn_reduce(const double *part, int nsplit, int M1, int M2, double alpha,
                            const double *sa, const double *sb, double beta, double *C, long ldc) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)M1 * M2) return;
  const int i = (int)(t % M1), j = (int)(t / M1);
  double s = 0.0;
  for (int q = 0; q < nsplit; q++) s += part[(long)q * M1 * M2 + t];
  const double v = alpha * s / (sa[i] * sb[j]);
  double *c = C + i + ldc * (long)j;
  *c = beta == 0.0 ? v : v + beta * *c;
}

// ----------------------------------------------------------------------------- NN
// pre-split C'' (hi, lo) in [j][k] layout, K padded to a multiple of 16, M2 padded to 32*...
__global__ void k_nn_prep(const double *C, long ldc, int K, int M2, const double *sx,
                          const double *tc, int Kp, _Float16 *ch, _Float16 *cl) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int M2p = (M2 + 127) / 128 * 128;
  if (t >= (long)Kp * M2p) return;
  const int k = (int)(t % Kp), j = (int)(t / Kp);
  double x = 0.0;
  if (k < K && j < M2) x = C[k + ldc * (long)j] / sx[k] * tc[j];
  const _Float16 h = to_half(x);
  ch[t] = h;
  cl[t] = to_half(x - (double)h);
}

// column maxima of C_kj / sx_k -> tc_j
__global__ __launch_bounds__(256) void k_nn_cscale(const double *C, long ldc, int K, int M2,
                                                   const double *sx, double *tc) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double m = 0.0;
  for (int k = threadIdx.x; k < K; k += 256) m = fmax(m, fabs(C[k + ldc * (long)j] / sx[k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double sc = 1.0;
    if (red[0] > 0.0) {
      int e;
      frexp(red[0], &e);
      sc = ldexp(1.0, 14 - e);
    }
    tc[j] = sc;
  }
}

struct NNArgs {
  long R;
  int K, M2, Kp;
  const double *X;
  long ldx;
  const double *sx;
  const _Float16 *ch, *cl;
  const double *tc;
  double alpha, beta;
  double *Y;
  long ldy;
  int row_blocks, col_tiles;
};

// Work-group: 8 waves; waves 0-3 cover rows [32w,32w+32) of a 128-row block, waves 4-7 the
// same rows for the second 128-column half: tile 128 rows x 256 columns, 4 blocks/wave.
template <int NT>
__global__ __launch_bounds__(512) void k_gemm_nn(NNArgs p) {
  const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
  // the col_tiles of one row block run consecutively on one XCD
  const long rb = (long)(jj / p.col_tiles) * 8 + xcd;
  const int ct = jj % p.col_tiles;
  if (rb >= p.row_blocks) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long r0 = rb * 128 + 32 * (w & 3);
  const int c0 = ct * 256 + 128 * (w >> 2);
  const long row = r0 + (lane & 31);
  const bool rin = row < p.R;
  f16v acc[4];
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[b][r] = 0.f;
  const int kh = 8 * (lane >> 5);
  for (int k0 = 0; k0 < p.K; k0 += 16) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = k0 + kh + q;
      v[q] = (rin && k < p.K) ? p.X[row + p.ldx * (long)k] * p.sx[k] : 0.0;
    }
    half8 ah, al;
    split8(v, 1.0, ah, al);
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int col = c0 + 32 * b + (lane & 31);
      const long off = (long)col * p.Kp + k0 + kh;
      const half8 bh = *(const half8 *)&p.ch[off];
      const half8 bl = *(const half8 *)&p.cl[off];
      acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[b], 0, 0, 0);
      acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[b], 0, 0, 0);
      if (NT == 4) acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bl, acc[b], 0, 0, 0);
    }
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int n = c0 + 32 * b + (lane & 31);
    const double f = n < p.M2 ? p.alpha / p.tc[n] : 0.0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const long m = r0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (m < p.R && n < p.M2) {
        const double v = f * (double)acc[b][r];
        double *y = p.Y + m + p.ldy * (long)n;
        *y = p.beta == 0.0 ? v : v + p.beta * *y;
      }
    }
  }
}

}  // namespace

void colscale(const double *A, long lda, long rows, int cols, const double *pre, double *scale,
              hipStream_t s) {
  if (cols == 0) return;
  Timed t("gemm_scale", 8.0 * (double)rows * cols);
  hipLaunchKernelGGL(k_colscale, dim3(cols), dim3(256), 0, s, A, lda, rows, cols, pre, scale);
  HIPCK(hipGetLastError());
}

void gemm_mixed_tn(int M1, int M2, long K, double alpha, const double *A, long lda,
                   const double *sa, const double *B, long ldb, const double *sb, double beta,
                   double *C, long ldc, int nterms, hipStream_t s) {
  if (M1 <= 0 || M2 <= 0) return;
  TNArgs p{};
  p.M1 = M1; p.M2 = M2; p.K = K; p.A = A; p.B = B; p.lda = lda; p.ldb = ldb; p.sa = sa; p.sb = sb;
  p.tiles_i = (M1 + TN_T - 1) / TN_T;
  p.tiles_j = (M2 + TN_T - 1) / TN_T;
  const int tiles = p.tiles_i * p.tiles_j;
  int nsplit = std::max(1, 2048 / tiles);
  const long kmin = 4096;  // at least 16 FP32 chunks of work per group
  nsplit = (int)std::max<long>(1, std::min<long>(nsplit, (K + kmin - 1) / kmin));
  long kper = (K + nsplit - 1) / nsplit;
  kper = (kper + TN_KS - 1) / TN_KS * TN_KS;
  nsplit = (int)((K + kper - 1) / kper);
  p.kper = kper;
  p.nsplit = nsplit;
  p.nsplit_pad = (nsplit + 7) / 8 * 8;
  p.part = (double *)scratch("gemm.tnpart", sizeof(double) * (size_t)nsplit * M1 * M2);
  {
    Timed t("gemm_mixed", 2.0 * (nterms >= 4 ? 4 : 3) * (double)M1 * M2 * K);  // f16 MFMA flops
    const unsigned grid = (unsigned)(p.nsplit_pad * tiles);
    if (nterms >= 4) hipLaunchKernelGGL(k_gemm_tn<4>, dim3(grid), dim3(256), 0, s, p);
    else hipLaunchKernelGGL(k_gemm_tn<3>, dim3(grid), dim3(256), 0, s, p);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_tn_reduce, dim3(nblk((long)M1 * M2, 256)), dim3(256), 0, s, p.part, nsplit,
                       M1, M2, alpha, sa, sb, beta, C, ldc);
    HIPCK(hipGetLastError());
  }
}

void gemm_mixed_nn(long R, int K, int M2, double alpha, const double *X, long ldx,
                   const double *sx, const double *C, long ldc, double beta, double *Y, long ldy,
                   int nterms, hipStream_t s) {
  if (R <= 0 || M2 <= 0) return;
  const int Kp = (K + 15) / 16 * 16;
  const int M2p = (M2 + 127) / 128 * 128;
  const int M2q = (M2 + 255) / 256 * 256;
  _Float16 *ch = (_Float16 *)scratch("gemm.nnch", sizeof(_Float16) * (size_t)Kp * M2q);
  _Float16 *cl = (_Float16 *)scratch("gemm.nncl", sizeof(_Float16) * (size_t)Kp * M2q);
  double *tc = (double *)scratch("gemm.nntc", sizeof(double) * M2q);
  hipLaunchKernelGGL(k_nn_cscale, dim3(M2), dim3(256), 0, s, C, ldc, K, M2, sx, tc);
  hipLaunchKernelGGL(k_nn_prep, dim3(nblk((long)Kp * M2p, 256)), dim3(256), 0, s, C, ldc, K, M2, sx,
                     tc, Kp, ch, cl);
  if (M2q > M2p) {  // zero the padding columns read by the last 256-wide tile
    HIPCK(hipMemsetAsync(ch + (size_t)Kp * M2p, 0, sizeof(_Float16) * (size_t)Kp * (M2q - M2p), s));
    HIPCK(hipMemsetAsync(cl + (size_t)Kp * M2p, 0, sizeof(_Float16) * (size_t)Kp * (M2q - M2p), s));
  }
  HIPCK(hipGetLastError());
  NNArgs p{};
  p.R = R; p.K = K; p.M2 = M2; p.Kp = Kp; p.X = X; p.ldx = ldx; p.sx = sx; p.ch = ch; p.cl = cl;
  p.tc = tc; p.alpha = alpha; p.beta = beta; p.Y = Y; p.ldy = ldy;
  p.row_blocks = (int)((R + 127) / 128);
  p.col_tiles = M2q / 256;
  const long groups = (p.row_blocks + 7) / 8;
  const unsigned grid = (unsigned)(groups * 8 * p.col_tiles);
  Timed t("gemm_mixed", 2.0 * (nterms >= 4 ? 4 : 3) * (double)R * M2 * K);  // f16 MFMA flops
  if (nterms >= 4) hipLaunchKernelGGL(k_gemm_nn<4>, dim3(grid), dim3(512), 0, s, p);
  else hipLaunchKernelGGL(k_gemm_nn<3>, dim3(grid), dim3(512), 0, s, p);
  HIPCK(hipGetLastError());
}

}  // namespace kukan
