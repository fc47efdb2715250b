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
//        C small.  X diag(sx) and C'' are pre-split into "fragment" layout: a fragment is
//        the 1 KB an MFMA operand register quad holds for 32 rows (columns) x 16 k, stored
//        lane by lane.  The main loop copies fragments HBM -> LDS with the gfx950 LDS-DMA
//        load (global_load_lds_dwordx4: lane i's 16 bytes land at base + 16 i), so staging
//        costs no registers and no LDS stores, and every LDS read is one conflict-free
//        ds_read_b128 per lane.
#include "gemm_mixed.hpp"
#include "kukan_internal.hpp"

#include <algorithm>
#include <cmath>

namespace kukan {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

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


// Per-column power-of-two scale s_j = 2^(14 - e_j) with max_i |A(i,j)| < 2^e_j (1 for a zero
// column).  Pass 1: grid (column, row chunk), each work-group reduces its chunk and folds it
// into the column maximum with a 64-bit atomic max on the bit pattern (non-negative doubles
// order like their bit patterns).  Pass 2 turns maxima into scales.
constexpr long CS_ROWS = 16384;  // rows per work-group
__global__ __launch_bounds__(256) void k_colmax(const double *A, long lda, long rows,
                                                unsigned long long *cmax) {
  __shared__ double red[256];
  const int j = blockIdx.y;
  const long r0 = (long)blockIdx.x * CS_ROWS, r1 = min(rows, r0 + CS_ROWS);
  const double *c = A + (long)j * lda;
  double m = 0.0;
  for (long i = r0 + threadIdx.x; i < r1; i += 256) m = fmax(m, fabs(c[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] > 0.0)
    atomicMax(cmax + j, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_maxscale(int cols, const unsigned long long *cmax, double *scale) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const double mx = __longlong_as_double((long long)cmax[j]);
  double sc = 1.0;
  if (mx > 0.0) {
    int e;
    frexp(mx, &e);  // mx < 2^e
    sc = ldexp(1.0, 14 - e);
  }
  scale[j] = sc;
}

// Scale from a norm: dV sum_x psi(x)^2 = g bounds |psi(x)| <= sqrt(g / dV) < 2^e, so
// 2^(13 - e) keeps |psi s| < 2^13 without a pass over psi (1 for g <= 0).
__global__ void k_normscale(int cols, const double *g, double dv, double *scale) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double sc = 1.0;
  if (g[j] > 0.0) {
    int e;
    frexp(sqrt(g[j] / dv), &e);
    sc = ldexp(1.0, 13 - e);
  }
  scale[j] = sc;
}

// ----------------------------------------------------------------------------- tiles
// Both products run on a 256 x 128 block tile, 8 waves of 64 x 64 (wave w: rows 64 (w & 3),
// columns 64 (w >> 2)), 32-deep K stages double-buffered in LDS as split halves
// [row][k] with a 40-half pitch (conflict-free 16-byte reads for 8-lane groups).  The next
// stage's global loads are issued before the current stage's MFMAs, so one barrier per
// stage suffices.
constexpr int BM = 256, BN = 128, BK = 32, PITCH = 40;

struct Stage {
  _Float16 ah[BM * PITCH], al[BM * PITCH], bh[BN * PITCH], bl[BN * PITCH];
};

// FP32 split of a block of 8 scaled values: hi = f16(x), lo = f16(x - hi) with x rounded
// to FP32 first (its 2^-24 rounding is below the split's 2^-22).
__device__ inline void split8f(const double *v, double s, half8 &hi, half8 &lo) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    float f = (float)(v[j] * s);
    asm volatile("" : "+v"(f));
    const _Float16 h = (_Float16)f;
    hi[j] = h;
    lo[j] = (_Float16)(f - (float)h);
  }
}

template <int NT>
__device__ inline void mma_stage(const Stage &S, int lane, int wi, int wj, f16v (&acc)[2][2]) {
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const int ko = 16 * s + 8 * (lane >> 5);
    half8 ah[2], al[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const int row = 64 * wi + 32 * a + (lane & 31);
      ah[a] = *(const half8 *)&S.ah[row * PITCH + ko];
      al[a] = *(const half8 *)&S.al[row * PITCH + ko];
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int col = 64 * wj + 32 * b + (lane & 31);
      const half8 bh = *(const half8 *)&S.bh[col * PITCH + ko];
      const half8 bl = *(const half8 *)&S.bl[col * PITCH + ko];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bh, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bh, acc[a][b], 0, 0, 0);
        if (NT == 4)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bl, acc[a][b], 0, 0, 0);
      }
    }
  }
}

// ----------------------------------------------------------------------------- TN
// Split-K: work-group (split q, tile) accumulates K-range [q kper, (q+1) kper) in FP32
// (kper <= 4096: <= 768 MFMA accumulations) and writes an FP32 partial tile; k_tn_reduce sums
// the partials in FP64 in a fixed order.  The tiles of one K-range are dispatched
// consecutively on one XCD (blockIdx % 8), so A and B stages are fetched from HBM once and
// re-read from that XCD's L2 by the other tiles.
struct TNArgs {
  int M1, M2;
  long K;
  const double *A, *B;
  long lda, ldb;
  const double *sa, *sb;
  float *part;  // [split][M1][M2], M2 fastest
  long kper;    // multiple of BK
  int nsplit, tiles_i, tiles_j, nsplit_pad;
};

template <int NT>
__global__ __launch_bounds__(512) void k_gemm_tn(TNArgs p) {
  __shared__ Stage S[2];
  const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
  const int tiles = p.tiles_i * p.tiles_j;
  const int split = (jj / tiles) * 8 + xcd;
  if (split >= p.nsplit) return;
  const int tile = jj % tiles;
  const int i0 = (tile % p.tiles_i) * BM, j0 = (tile / p.tiles_i) * BN;
  const long kbeg = (long)split * p.kper, kend = min(p.K, kbeg + p.kper);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wi = w & 3, wj = w >> 2;
  // staging: A units (column t/4 + 128 v, k-group t%4), B unit (column t/4, k-group t%4)
  // Staging map: thread t holds the k pair 2 (t & 15), 2 (t & 15) + 1 of columns (t >> 4) + 32 v
  // (A: v < 8, B: v < 4), loaded as one 16-byte access, so 16 lanes read one column's 256-byte
  // stage segment and a wave instruction covers 4 whole segments.  gload only issues loads
  // (columns clamped into the operand, so no load is guarded); masking, scaling and
  // splitting happen in lstore, after the MFMAs of the current stage.
  const int kp = 2 * (t & 15), cg = t >> 4;
  const double *colA[8], *colB[4];
  double sav[8], sbv[4];
#pragma unroll
  for (int v = 0; v < 8; v++) {
    const int ia = i0 + cg + 32 * v;
    sav[v] = ia < p.M1 ? p.sa[ia] : 0.0;
    colA[v] = p.A + (long)min(ia, p.M1 - 1) * p.lda;
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int jb = j0 + cg + 32 * v;
    sbv[v] = jb < p.M2 ? p.sb[jb] : 0.0;
    colB[v] = p.B + (long)min(jb, p.M2 - 1) * p.ldb;
  }
  typedef double d2 __attribute__((ext_vector_type(2)));
  d2 ra[8], rb[4];
  long k0_cur = 0;
  auto gload = [&](long k0) {
    k0_cur = k0;
    const long kk = k0 + kp;
    if (k0 + BK <= kend) {  // whole stage inside the K-range (uniform across the block)
#pragma unroll
      for (int v = 0; v < 8; v++) ra[v] = *(const d2 *)(colA[v] + kk);
#pragma unroll
      for (int v = 0; v < 4; v++) rb[v] = *(const d2 *)(colB[v] + kk);
    } else {
      const long k1 = min(kk + 1, kend - 1), kc = min(kk, kend - 1);
#pragma unroll
      for (int v = 0; v < 8; v++) ra[v] = d2{colA[v][kc], colA[v][k1]};
#pragma unroll
      for (int v = 0; v < 4; v++) rb[v] = d2{colB[v][kc], colB[v][k1]};
    }
  };
  typedef _Float16 half2v __attribute__((ext_vector_type(2)));
  auto put2 = [&](_Float16 *hi, _Float16 *lo, int row, d2 x, double sc, bool ok0, bool ok1) {
    float f0 = ok0 ? (float)(x[0] * sc) : 0.f, f1 = ok1 ? (float)(x[1] * sc) : 0.f;
    asm volatile("" : "+v"(f0), "+v"(f1));
    const _Float16 h0 = (_Float16)f0, h1 = (_Float16)f1;
    *(half2v *)&hi[row * PITCH + kp] = half2v{h0, h1};
    *(half2v *)&lo[row * PITCH + kp] = half2v{(_Float16)(f0 - (float)h0), (_Float16)(f1 - (float)h1)};
  };
  auto lstore = [&](Stage &D) {
    const bool ok0 = k0_cur + kp < kend, ok1 = k0_cur + kp + 1 < kend;
#pragma unroll
    for (int v = 0; v < 8; v++) put2(D.ah, D.al, cg + 32 * v, ra[v], sav[v], ok0, ok1);
#pragma unroll
    for (int v = 0; v < 4; v++) put2(D.bh, D.bl, cg + 32 * v, rb[v], sbv[v], ok0, ok1);
  };
  f16v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
  const int nst = (int)((kend - kbeg + BK - 1) / BK);
  gload(kbeg);
  lstore(S[0]);
  __syncthreads();
  for (int st = 0; st < nst; st++) {
    const bool more = st + 1 < nst;
    if (more) gload(kbeg + (long)(st + 1) * BK);
    mma_stage<NT>(S[st & 1], lane, wi, wj, acc);
    if (more) lstore(S[(st + 1) & 1]);
    __syncthreads();
  }
  float *out = p.part + (long)split * p.M1 * p.M2;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = i0 + 64 * wi + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = j0 + 64 * wj + 32 * b + (lane & 31);
        if (m < p.M1 && n < p.M2) out[(long)m * p.M2 + n] = acc[a][b][r];
      }
}

// tri = 1 / 2: only the 256 x 256 tiles on and below / above the diagonal were computed;
// the others are left as they were
__global__ void k_tn_reduce(const float *part, int nsplit, int M1, int M2, double alpha,
                            const double *sa, const double *sb, double beta, double *C, long ldc,
                            int tri) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)M1 * M2) return;
  const int i = (int)(t / M2), j = (int)(t % M2);
  if ((tri == 1 && (i >> 8) < (j >> 8)) || (tri == 2 && (i >> 8) > (j >> 8))) return;
  // in split order (unrolled so that several partial loads are in flight)
  double s = 0.0;
#pragma unroll 16
  for (int q = 0; q < nsplit; q++) s += (double)part[(long)q * M1 * M2 + t];
  const double v = alpha * s / (sa[i] * sb[j]);
  double *c = C + i + ldc * (long)j;
  *c = beta == 0.0 ? v : v + beta * *c;
}

// ----------------------------------------------------------------------------- NN
// C'' = C / sx * tc split into hi/lo halves in fragment layout: fragment (j / 32, k / 16) of
// KC = Kp / 16 per column block, lane (j % 32) + 32 ((k % 16) / 8), half k % 8.  Columns
// are padded to M2p and k to Kp with zeros.
__global__ void k_nn_prep(const double *C, long ldc, int K, int M2, const double *sx,
                          const double *tc, int Kp, int M2p, _Float16 *ch, _Float16 *cl) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;  // = fragment element index
  if (t >= (long)Kp * M2p) return;
  const int hlf = (int)(t & 7), lane = (int)((t >> 3) & 63);
  const long frag = t >> 9;
  const int KC = Kp / 16;
  const int j = (int)(frag / KC) * 32 + (lane & 31);
  const int k = (int)(frag % KC) * 16 + 8 * (lane >> 5) + hlf;
  double x = 0.0;
  if (k < K && j < M2) x = C[k + ldc * (long)j] / sx[k] * tc[j];
  const _Float16 h = to_half(x);
  ch[t] = h;
  cl[t] = to_half(x - (double)h);
}

// column maxima of C_kj / sx_k -> tc_j
__global__ __launch_bounds__(256) void k_nn_cscale(const double *C, long ldc, int K, int M2,
                                                   const double *sx, double alpha, double *tc,
                                                   double *ys) {
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
    ys[j] = alpha / sc;
  }
}

// X'' = X diag(sx) split into hi/lo halves in fragment layout: fragment (row / 32, k / 16)
// at ((row / 32) KC + k / 16) 512 halves (KC = Kp / 16), lane (row % 32) + 32 ((k % 16) / 8),
// half k % 8; rows padded to a multiple of NN_BM and k to Kp with zeros.  Work-group: 64
// rows x 64 k through LDS; global reads are 64-row runs of one column, writes are whole
// fragments (lane-contiguous 16-byte runs).
constexpr int NN_BM = 256, NN_BN = 256;
__global__ __launch_bounds__(256) void k_split_t(const double *X, long ldx, long R, int K,
                                                 const double *sx, int Kp, _Float16 *xh,
                                                 _Float16 *xl) {
  __shared__ float tile[64][65];
  const long r0 = (long)blockIdx.x * 64;
  const int k0 = blockIdx.y * 64;
  const int t = threadIdx.x, tr = t & 63, tk = t >> 6;
  const long r = r0 + tr;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int k = k0 + tk + 4 * q;
    float v = 0.f;
    if (r < R && k < K) v = (float)(X[r + ldx * (long)k] * sx[k]);
    tile[tk + 4 * q][tr] = v;
  }
  __syncthreads();
  const int KC = Kp / 16;
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int unit = t + 256 * u;  // 8 fragments (2 row blocks x 4 k blocks) x 64 lanes
    const int lane = unit & 63, fr = unit >> 6, rb = fr >> 2, kc = fr & 3;
    const int rl = 32 * rb + (lane & 31), kl = 16 * kc + 8 * (lane >> 5);
    half8 h, l;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float f = tile[kl + j][rl];
      asm volatile("" : "+v"(f));
      const _Float16 hh = (_Float16)f;
      h[j] = hh;
      l[j] = (_Float16)(f - (float)hh);
    }
    const long off = (((r0 / 32 + rb) * KC) + k0 / 16 + kc) * 512 + 8 * lane;
    *(half8 *)&xh[off] = h;
    *(half8 *)&xl[off] = l;
  }
}

struct NNArgs {
  long R;
  int M2, Kp;
  const _Float16 *xh, *xl;  // X'' fragments, (Rp / 32) x (Kp / 16)
  const _Float16 *ch, *cl;  // C'' fragments, (M2p / 32) x (Kp / 16)
  const double *ys;  // alpha / tc per column of Y
  double beta;
  double *Y;
  long ldy;
  int row_blocks, col_tiles;
  double *colsq;  // optional (beta = 0): sums of Y^2 per column and row block, [col][row block]
  int cupper;     // C upper triangular: column tile ct needs only k < 256 (ct + 1)
};

// s_waitcnt with only the vector-memory count constrained (expcnt, lgkmcnt at their maxima)
template <int N>
__device__ inline void wait_vm() {
  static_assert(N >= 0 && N < 16, "vmcnt immediate");
  __builtin_amdgcn_s_waitcnt(0xF70 | N);
}

#define NN_LDS(x) ((__attribute__((address_space(3))) void *)(x))

// LDS written by lanes of this wave, next read by other lanes of the same wave (no other wave
// touches the area): a wavefront-scope fence pair around the wave barrier
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int NN_NS = 5;
// ring of NS stages for a work-group of 2 WR waves over 64 WR rows x 256 columns: per stage
// 2 WR A fragments (hi), 2 WR (lo), then 8 B fragments (hi) and 8 (lo), 1 KB each
template <int WR, int NS>
using FragRing = _Float16[NS][4 * WR + 16][512];

// wait until this wave's loads of all but the `later` most recent stages (LPS loads each) have
// landed; later <= MAXL
template <int LPS, int MAXL>
__device__ inline void wait_stages(int later) {
  static_assert(LPS * MAXL < 16, "vmcnt immediate");
  if constexpr (MAXL >= 3) {
    if (later >= 3) { wait_vm<3 * LPS>(); return; }
  }
  if constexpr (MAXL >= 2) {
    if (later >= 2) { wait_vm<2 * LPS>(); return; }
  }
  if constexpr (MAXL >= 1) {
    if (later >= 1) { wait_vm<LPS>(); return; }
  }
  wait_vm<0>();
}

// The fragment-GEMM main loop shared by NN and pre-split TN, for a work-group of 2 WR waves:
// acc (wave w: rows 64 (w % WR), columns 128 (w / WR) of a 64 WR x 256 tile) += the products
// of nst 16-deep stages.  The tile's A fragments of stage st are at a[h|l] + (idx KC + st) 512
// (idx = row block 0 .. 2 WR - 1), the B fragments likewise (idx 0..7); the caller offsets
// a* / b* to its first row / column block and first stage.  Each wave issues LPS = 2 + 8 / WR
// LDS-DMA loads per stage (buffer_load ... lds through one buffer descriptor per fragment row:
// the MUBUF form, unlike the global one, leaves the compiler's LDS-read waits graded), NS - 1
// stages ahead of the one it computes.  WR = 4, NS = 5 (160 KB, one work-group per CU) and
// WR = 2, NS = 3 (72 KB, two per CU) are used.  Stage st + 1's
// fragments are read from LDS while stage st multiplies: the next A fragments after the first
// MFMA group, each B fragment as soon as its last MFMA of stage st is issued (at most 12 LDS
// reads outstanding, within the 4-bit lgkm counter).  Before reading stage st + 1 a wave waits
// for its own loads of that stage (vmcnt counts them in issue order) and a barrier publishes
// everyone's; the buffer refilled after that barrier held stage st - 1, which every wave had
// read into registers and consumed before arriving.  The plain s_barrier (no fence) keeps the
// later stages in flight; empty asm statements keep the compiler from moving LDS accesses
// across it.  The accumulation order per output (hi hi, hi lo, [lo lo,] lo hi) is the same in
// every stage.  On return this wave's loads have landed; the block is not synchronised: the
// last barrier was at the start of stage nst - 2.  Slower waves may still read the ring slot
// of stage nst - 1 (its fragments are read during stage nst - 2), and their reads of stage
// nst - 2's slot, issued during stage nst - 3, may still be outstanding when they pass that
// barrier.  The slots of stages nst - 3 and earlier were read into registers and consumed by
// MFMAs before every wave reached the barrier, so NS - 2 slots are free.
// Fragment rows are addressed through 32-bit buffer offsets: KC 16-deep stages of 1 KB each.
static void frag_rows_ok(long KC, const char *who) {
  if (KC * 1024 > 0x7fffffffL) throw Error(1, std::string(who) + ": operand too long for the split-precision product");
}

// ring slot of stage st: NF fragments of 512 halves
template <int NF, int NS>
__device__ inline _Float16 (*frag_slot(_Float16 *ring, int st))[512] {
  return reinterpret_cast<_Float16(*)[512]>(ring + (long)(st % NS) * NF * 512);
}

// this wave's LPS LDS-DMA loads of stage st (fragments LPS ws .. LPS ws + LPS - 1)
template <int LPS, int NF, int NS>
__device__ inline void frag_dma(_Float16 *ring, const __amdgpu_buffer_rsrc_t *rs, int ws, int lane,
                                int st) {
  _Float16(*d)[512] = frag_slot<NF, NS>(ring, st);
#pragma unroll
  for (int i = 0; i < LPS; i++)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs[i], NN_LDS(&d[LPS * ws + i][0]), 16, st * 1024 + 16 * lane, 0, 0, 0);
}

template <int NT, int WR, int NS>
__device__ inline void frag_gemm(const _Float16 *ah_g, const _Float16 *al_g, const _Float16 *bh_g,
                                 const _Float16 *bl_g, long KC, int nst, _Float16 *ring,
                                 f16v (&acc)[2][4]) {
  static_assert((WR == 2 || WR == 4) && NS >= 3, "ring shape");
  constexpr int NA = 2 * WR, LPS = 2 + 8 / WR;  // A fragments per half; loads per wave and stage
  constexpr int NF = 2 * NA + 16;               // fragments per stage (ring slot)
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wi = w % WR, wj = w / WR;
  // this wave's LPS fragments per stage: f = LPS w + i; f < NA: A hi, < 2 NA: A lo, then 8
  // B hi and 8 B lo.  One buffer descriptor per fragment row (uniform per wave); a row spans
  // KC * 1 KB, which the host keeps below 2^31 bytes (frag_rows_ok)
  const int ws = __builtin_amdgcn_readfirstlane(w);
  __amdgpu_buffer_rsrc_t rs[LPS];
#pragma unroll
  for (int i = 0; i < LPS; i++) {
    const int f = LPS * ws + i;
    const _Float16 *base;
    int idx;
    if (f < NA) { base = ah_g; idx = f; }
    else if (f < 2 * NA) { base = al_g; idx = f - NA; }
    else if (f < 2 * NA + 8) { base = bh_g; idx = f - 2 * NA; }
    else { base = bl_g; idx = f - 2 * NA - 8; }
    rs[i] = __builtin_amdgcn_make_buffer_rsrc((void *)(base + (long)idx * KC * 512), (short)0,
                                              (int)(KC * 1024), 0x00020000);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[a][b][r] = 0.f;
#pragma unroll
  for (int i = 0; i < NS - 1; i++)
    if (i < nst) frag_dma<LPS, NF, NS>(ring, rs, ws, lane, i);
  wait_stages<LPS, NS - 2>(min(NS - 2, nst - 1));
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  half8 ah[2], al[2], bh[4], bl[4];
  {
    const _Float16(*C)[512] = frag_slot<NF, NS>(ring, 0);
#pragma unroll
    for (int a = 0; a < 2; a++) {
      ah[a] = *(const half8 *)&C[2 * wi + a][8 * lane];
      al[a] = *(const half8 *)&C[NA + 2 * wi + a][8 * lane];
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
      bh[b] = *(const half8 *)&C[2 * NA + 4 * wj + b][8 * lane];
      bl[b] = *(const half8 *)&C[2 * NA + 8 + 4 * wj + b][8 * lane];
    }
  }
  auto stage = [&](auto more_t, int st) {
    constexpr bool more = decltype(more_t)::value;
    half8 ahn[2], aln[2];
    const _Float16(*C1)[512] = frag_slot<NF, NS>(ring, st + 1);
    if (more) {
      // stage st + 1 landed (own loads) and published; stages after it stay in flight
      wait_stages<LPS, NS - 3>(min(NS - 3, nst - 2 - st));
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      // the buffer refilled held stage st - 1, whose fragments every wave has consumed
      if (st + NS - 1 < nst) frag_dma<LPS, NF, NS>(ring, rs, ws, lane, st + NS - 1);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int b = 0; b < 4; b++) {
#pragma unroll
      for (int a = 0; a < 2; a++) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 2; a++) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
      if (NT == 4) {
#pragma unroll
        for (int a = 0; a < 2; a++) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bl[b], acc[a][b], 0, 0, 0);
      }
      if (more && b == 0) {
#pragma unroll
        for (int a = 0; a < 2; a++) {
          ahn[a] = *(const half8 *)&C1[2 * wi + a][8 * lane];
          aln[a] = *(const half8 *)&C1[NA + 2 * wi + a][8 * lane];
        }
      }
      if (more) bl[b] = *(const half8 *)&C1[2 * NA + 8 + 4 * wj + b][8 * lane];
#pragma unroll
      for (int a = 0; a < 2; a++) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
      if (more) bh[b] = *(const half8 *)&C1[2 * NA + 4 * wj + b][8 * lane];
      __builtin_amdgcn_sched_barrier(0);
    }
    if (more) {
#pragma unroll
      for (int a = 0; a < 2; a++) {
        ah[a] = ahn[a];
        al[a] = aln[a];
      }
    }
  };
  for (int st = 0; st + 1 < nst; st++) stage(std::true_type{}, st);
  if (nst > 0) stage(std::false_type{}, nst - 1);
  wait_vm<0>();
}

// Y tile 64 WR rows x 256 columns (frag_gemm over all of K): WR = 4 (256 rows, 160 KB ring,
// one work-group per CU) or WR = 2 (128 rows, 72 KB ring, two work-groups per CU, so one's
// epilogue and load waits overlap the other's MFMAs).  Every output gets the same MFMA
// sequence in both shapes, so Y is the same bit for bit.  The col_tiles of one row block run
// consecutively on one XCD so that X'' is fetched from HBM once.  The epilogue goes through
// LDS so that each column of Y is written (and, for beta != 0, read) as runs of 64 rows.
template <int NT, int WR, int NS>
__global__ __launch_bounds__(128 * WR, 2) void k_gemm_nn(NNArgs p) {
  constexpr int NF = 4 * WR + 16;  // fragments per ring slot
  __shared__ __attribute__((aligned(16))) _Float16 ring[NS * NF * 512];
  const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
  const long rb = (long)(jj / p.col_tiles) * 8 + xcd;
  const int ct = jj % p.col_tiles;
  if (rb >= p.row_blocks) return;
  const long r0 = rb * 64 * WR;
  const int c0 = ct * NN_BN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wi = w % WR, wj = w / WR;
  const long KC = p.Kp / 16;
  // with C upper triangular, rows k >= c0 + 256 of C are zero for these columns: their
  // products would add exact zeros, so the stages are skipped
  const int nst = p.cupper ? (int)min(KC, (long)(c0 + NN_BN) / 16) : (int)KC;
  f16v acc[2][4];
  frag_gemm<NT, WR, NS>(p.xh + (r0 / 32) * KC * 512, p.xl + (r0 / 32) * KC * 512,
                        p.ch + (long)(c0 / 32) * KC * 512, p.cl + (long)(c0 / 32) * KC * 512, KC,
                        nst, ring, acc);
  // epilogue: per 32-column block b, each wave parks its 64 x 32 piece in LDS as
  // [column][row] (pitch 65 floats: conflict-free) and writes it back column by column.
  // For beta != 0 the block's 32 Y values per lane are loaded before the parking, as
  // independent loads whose latency overlaps the LDS round trip.  Each wave's area (and its
  // column sums) is its own, so the waves synchronise only with themselves until the column
  // sums are combined.
  // WR = 4: the areas lie in the three ring slots of stages nst - 3, nst - 4 and nst - 5
  // (slots last + 3, + 2, + 1), which no wave reads after the last barrier (stage nst - 1's
  // slot may still be read, and stage nst - 2's may still have reads in flight): three
  // 8320-byte wave areas per 32 KB slot, waves 6 and 7 with the column sums behind them in the
  // third, so a wave starts its epilogue without waiting for the others.  WR = 2: the ring has
  // one free slot only; the work-group synchronises once and the areas start at the ring's
  // base (the other work-group on the CU keeps the SIMDs busy meanwhile).
  constexpr int EP = 65;
  float *W;
  double *cs;  // column sums of squares of the parked values, [row group wi][tile column]
  if constexpr (WR == 4) {
    constexpr size_t slot_bytes = NF * 512 * sizeof(_Float16);
    static_assert(NS == 5 && 3 * 32 * EP * sizeof(float) <= slot_bytes &&
                      2 * 32 * EP * sizeof(float) + 1024 * sizeof(double) <= slot_bytes,
                  "epilogue areas: three free ring slots, three wave areas per slot");
    const int last = (nst - 1) % NS;  // nst >= 1
    W = reinterpret_cast<float *>(ring + (long)((last + 1 + w / 3) % NS) * NF * 512) + (w % 3) * 32 * EP;
    // after the two wave areas of slot last + 3 (2 * 32 * EP floats = 16640 bytes, 8-byte
    // aligned)
    cs = reinterpret_cast<double *>(reinterpret_cast<float *>(ring + (long)((last + 3) % NS) * NF * 512) +
                                    2 * 32 * EP);
  } else {
    static_assert(2 * WR * 32 * EP * sizeof(float) + WR * 256 * sizeof(double) <= sizeof(ring),
                  "epilogue areas: the whole ring");
    __syncthreads();  // every wave's last fragment reads are done
    float *base = reinterpret_cast<float *>(ring);
    W = base + w * 32 * EP;
    cs = reinterpret_cast<double *>(base + 2 * WR * 32 * EP);  // 8-byte aligned
  }
  const long m = r0 + 64 * wi + lane;
  const bool mrow = m < p.R;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int nb0 = c0 + 128 * wj + 32 * b;
    double yo[32];
    if (p.beta != 0.0) {
#pragma unroll
      for (int q = 0; q < 32; q++)
        yo[q] = (mrow && nb0 + q < p.M2) ? p.Y[m + p.ldy * (long)(nb0 + q)] : 0.0;
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int ml = 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        W[(lane & 31) * EP + ml] = acc[a][b][r];
      }
    wave_lds_sync();
    if (p.colsq) {  // lane: column lane % 32, rows 32 (lane / 32) .. +31 of the wave's 64
      const int q = lane & 31, h = lane >> 5;
      double s2 = 0.0;
#pragma unroll 8
      for (int r = 0; r < 32; r++) {
        const double x = (double)W[q * EP + 32 * h + r];
        s2 += x * x;
      }
      s2 += __shfl_down(s2, 32, 64);
      if (lane < 32) cs[wi * 256 + 128 * wj + 32 * b + q] = s2;
    }
#pragma unroll
    for (int q = 0; q < 32; q++) {
      const int n = nb0 + q;
      if (n < p.M2 && mrow) {
        const double v = p.ys[n] * (double)W[q * EP + lane];
        p.Y[m + p.ldy * (long)n] = p.beta == 0.0 ? v : v + p.beta * yo[q];
      }
    }
    wave_lds_sync();  // the next block's parking overwrites this wave's area
  }
  if (p.colsq) __syncthreads();  // column sums of all waves
  if (p.colsq && t < 256 && c0 + t < p.M2) {  // the values written are ys W exactly
    const double y2 = p.ys[c0 + t] * p.ys[c0 + t];
    double sum = cs[t];
#pragma unroll
    for (int g = 1; g < WR; g++) sum += cs[g * 256 + t];
    p.colsq[(long)(c0 + t) * p.row_blocks + rb] = y2 * sum;
  }
}

// ----------------------------------------------------------------------------- TN, pre-split
// A (K x M) column-major times diag(s), split into hi/lo halves in fragment layout: fragment
// (i / 32, k / 16) at ((i / 32) KC + k / 16) 512 halves (KC = Kp / 16), lane (i % 32) +
// 32 ((k % 16) / 8), half k % 8; columns padded to Mp and k to Kp with zeros.  Work-group:
// 32 columns x 128 k through LDS.  Reads: 8 lanes cover one column's 128-byte run; writes:
// the 8 fragments of the tile are consecutive in memory (16 KB of hi and lo in all).
constexpr int SPT_K = 128;
__global__ __launch_bounds__(256) void k_split_tn(const double *A, long lda, long K, int M,
                                                  const double *sc, long KC, _Float16 *xh,
                                                  _Float16 *xl) {
  __shared__ float tile[32][SPT_K + 4];
  const int ib = blockIdx.y;
  const long k0 = (long)blockIdx.x * SPT_K;
  const int t = threadIdx.x;
  {
    const int cl = t >> 3, kq = 2 * (t & 7);  // column, k pair within a 16-k step
    const int i = ib * 32 + cl;
    const bool col = i < M;
    const double si = col ? sc[i] : 0.0;
    const double *a = A + (long)(col ? i : 0) * lda;
#pragma unroll
    for (int q = 0; q < SPT_K / 16; q++) {
      const long k = k0 + 16 * q + kq;
      const double v0 = (col && k < K) ? a[k] : 0.0, v1 = (col && k + 1 < K) ? a[k + 1] : 0.0;
      tile[cl][16 * q + kq] = (float)(v0 * si);
      tile[cl][16 * q + kq + 1] = (float)(v1 * si);
    }
  }
  __syncthreads();
  const long kc0 = k0 / 16;
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int slot = t + 256 * u;  // 8 fragments x 64 lanes
    const int lane = slot & 63, f = slot >> 6;
    const int cl = lane & 31, kl = 16 * f + 8 * (lane >> 5);
    half8 h, l;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float v = tile[cl][kl + j];
      asm volatile("" : "+v"(v));
      const _Float16 hh = (_Float16)v;
      h[j] = hh;
      l[j] = (_Float16)(v - (float)hh);
    }
    const long off = (((long)ib * KC + kc0 + f) * 64 + lane) * 8;
    *(half8 *)&xh[off] = h;
    *(half8 *)&xl[off] = l;
  }
}

// X diag(s) (R x M, column-major) split once into both fragment layouts: the TN layout
// (fragment (i / 32, k / 16) of X as the first operand of X^T B, KC = Rp / 16 per band block)
// and the NN layout (fragment (row / 32, band / 16) of X as the left factor of X C, KCn per
// row block).  Work-group: 64 rows x 64 bands through LDS; each thread writes two 16-byte lane
// slots of each layout.  Bands at or beyond 16 KCn have no NN fragments.
__global__ __launch_bounds__(256) void k_split_both(const double *X, long ldx, long R, int M,
                                                    const double *sc, long KC, int KCn,
                                                    _Float16 *th, _Float16 *tl, _Float16 *nh,
                                                    _Float16 *nl) {
  __shared__ float tile[64][65];  // [band][row]
  const long r0 = (long)blockIdx.x * 64;
  const int b0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  // loads: a wave reads one band's 64-row run (512 contiguous bytes) per instruction
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = q * 256 + t, bl = e >> 6, rl = e & 63;
    const int b = b0 + bl;
    const long r = r0 + rl;
    tile[bl][rl] = (b < M && r < R) ? (float)(X[(long)b * ldx + r] * sc[b]) : 0.f;
  }
  __syncthreads();
  auto split8 = [&](const float *v, half8 &h, half8 &l) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float f = v[j];
      asm volatile("" : "+v"(f));
      const _Float16 hh = (_Float16)f;
      h[j] = hh;
      l[j] = (_Float16)(f - (float)hh);
    }
  };
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int slot = t + 256 * u;  // 8 fragments x 64 lanes
    const int lane = slot & 63, fr = slot >> 6;
    // TN: band block fr >> 2 (of 2), row chunk fr & 3 (of 4)
    {
      const int bb = fr >> 2, kc = fr & 3;
      const int bl = 32 * bb + (lane & 31), rl = 16 * kc + 8 * (lane >> 5);
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) v[j] = tile[bl][rl + j];
      half8 h, l;
      split8(v, h, l);
      const long off = ((((long)b0 / 32 + bb) * KC + r0 / 16 + kc) * 64 + lane) * 8;
      *(half8 *)&th[off] = h;
      *(half8 *)&tl[off] = l;
    }
    // NN: row block fr >> 2 (of 2), band chunk fr & 3 (of 4)
    {
      const int rb = fr >> 2, bc = fr & 3;
      const int rl = 32 * rb + (lane & 31), bl = 16 * bc + 8 * (lane >> 5);
      const int bcg = b0 / 16 + bc;
      if (bcg < KCn) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = tile[bl + j][rl];
        half8 h, l;
        split8(v, h, l);
        const long off = (((r0 / 32 + rb) * KCn + bcg) * 64 + lane) * 8;
        *(half8 *)&nh[off] = h;
        *(half8 *)&nl[off] = l;
      }
    }
  }
}

struct TNFArgs {
  int M1, M2;
  const _Float16 *ah, *al, *bh, *bl;  // fragments, (M1p / 32) x KC and (M2p / 32) x KC
  long KC;
  float *part;  // [split][M1][M2], M2 fastest
  int kper_st;  // stages (16 k) per split
  int nsplit, tiles_i, tiles_j;
  int tri;  // 0: all tiles; 1 / 2: the tiles on and below / above the diagonal (M1 = M2)
};

// Split-K TN from fragments: work-group = one 256 x 256 output tile over one K-range of
// kper_st stages (FP32 accumulation depth 16 kper_st <= 4096), written as an FP32 partial
// tile for k_tn_reduce.  The tiles of one K-range run consecutively on one XCD so that its
// A and B fragments come from HBM once.
template <int NT>
__global__ __launch_bounds__(512) void k_gemm_tnf(TNFArgs p) {
  __shared__ __attribute__((aligned(16))) FragRing<4, NN_NS> buf;
  const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
  const int tiles = p.tri ? p.tiles_i * (p.tiles_i + 1) / 2 : p.tiles_i * p.tiles_j;
  const int split = (jj / tiles) * 8 + xcd;
  if (split >= p.nsplit) return;
  const int tile = jj % tiles;
  int ti = tile % p.tiles_i, tj = tile / p.tiles_i;
  if (p.tri) {  // tile = ti (ti + 1) / 2 + tj with tj <= ti (lower); transposed for upper
    ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= tile) ti++;
    tj = tile - ti * (ti + 1) / 2;
    if (p.tri == 2) {
      const int x = ti;
      ti = tj;
      tj = x;
    }
  }
  const int i0 = ti * 256, j0 = tj * 256;
  const long st0 = (long)split * p.kper_st;
  const int nst = (int)min((long)p.kper_st, p.KC - st0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wi = w & 3, wj = w >> 2;
  f16v acc[2][4];
  const long ao = ((long)(i0 / 32) * p.KC + st0) * 512, bo = ((long)(j0 / 32) * p.KC + st0) * 512;
  frag_gemm<NT, 4, NN_NS>(p.ah + ao, p.al + ao, p.bh + bo, p.bl + bo, p.KC, nst, &buf[0][0][0], acc);
  float *out = p.part + (long)split * p.M1 * p.M2;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = i0 + 64 * wi + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = j0 + 128 * wj + 32 * b + (lane & 31);
        if (m < p.M1 && n < p.M2) out[(long)m * p.M2 + n] = acc[a][b][r];
      }
}

}  // namespace

void colscale(const double *A, long lda, long rows, int cols, double *scale, hipStream_t s) {
  if (cols <= 0) return;
  if (rows <= 0) throw Error(1, "colscale: empty columns");
  Timed t("gemm_scale", 8.0 * (double)rows * cols);
  auto *cmax = (unsigned long long *)scratch("gemm.cmax", sizeof(unsigned long long) * cols);
  launch_fill(cols, 0.0, (double *)cmax, s);  // all-zero bits
  const long chunks = (rows + CS_ROWS - 1) / CS_ROWS;
  hipLaunchKernelGGL(k_colmax, dim3((unsigned)chunks, (unsigned)cols), dim3(256), 0, s, A, lda, rows,
                     cmax);
  hipLaunchKernelGGL(k_maxscale, dim3(nblk(cols, 256)), dim3(256), 0, s, cols, cmax, scale);
  HIPCK(hipGetLastError());
}

// TN from pre-split fragments: A diag(sa) and B diag(sb) are split once each (one pass when
// B is A with the same scales), then k_gemm_tnf and k_tn_reduce.
static void gemm_mixed_tn_presplit(int M1, int M2, long K, double alpha, const double *A, long lda,
                                   const double *sa, const double *B, long ldb, const double *sb,
                                   double beta, double *C, long ldc, int nterms, hipStream_t s,
                                   const SplitBands *pa, int tri) {
  if (tri && M1 != M2) throw Error(1, "gemm_mixed_tn: a triangle needs a square product");
  const long Kp = (K + 255) / 256 * 256, KC = Kp / 16;  // whole split tiles (as split_bands)
  frag_rows_ok(KC, "gemm_mixed_tn");
  const int M1p = (M1 + 255) / 256 * 256, M2p = (M2 + 255) / 256 * 256;
  const bool same = A == B && sa == sb && lda == ldb && M1 == M2;
  if (pa && !(pa->X == A && pa->ldx == lda && pa->R == K && pa->M == M1 && pa->s == sa))
    throw Error(1, "gemm_mixed_tn: pre-split operand does not match A");
  _Float16 *ah = pa ? pa->th : (_Float16 *)scratch("gemm.tnah", sizeof(_Float16) * (size_t)M1p * Kp);
  _Float16 *al = pa ? pa->tl : (_Float16 *)scratch("gemm.tnal", sizeof(_Float16) * (size_t)M1p * Kp);
  _Float16 *bh = ah, *bl = al;
  {
    Timed t("gemm_split", (8.0 * K + 4.0 * Kp) * ((pa ? 0 : M1) + (same ? 0 : M2)));
    if (!pa)
      hipLaunchKernelGGL(k_split_tn, dim3((unsigned)(Kp / SPT_K), (unsigned)(M1p / 32)), dim3(256), 0,
                         s, A, lda, K, M1, sa, KC, ah, al);
    if (!same) {
      bh = (_Float16 *)scratch("gemm.tnbh", sizeof(_Float16) * (size_t)M2p * Kp);
      bl = (_Float16 *)scratch("gemm.tnbl", sizeof(_Float16) * (size_t)M2p * Kp);
      hipLaunchKernelGGL(k_split_tn, dim3((unsigned)(Kp / SPT_K), (unsigned)(M2p / 32)), dim3(256), 0,
                         s, B, ldb, K, M2, sb, KC, bh, bl);
    }
    HIPCK(hipGetLastError());
  }
  TNFArgs p{};
  p.M1 = M1; p.M2 = M2; p.ah = ah; p.al = al; p.bh = bh; p.bl = bl; p.KC = KC;
  p.tiles_i = M1p / 256;
  p.tiles_j = M2p / 256;
  p.tri = tri;
  // stages per K-range: 256 (4096 k, FP32 accumulation depth), more if the FP32 partials
  // would exceed 2^29 values (2 GiB)
  const long max_split = std::max<long>(1, (1L << 29) / ((long)M1 * M2));
  p.kper_st = (int)std::max<long>(256, (KC + max_split - 1) / max_split);
  p.nsplit = (int)((KC + p.kper_st - 1) / p.kper_st);
  p.part = (float *)scratch("gemm.tnpart", sizeof(float) * (size_t)p.nsplit * M1 * M2);
  const int tiles = tri ? p.tiles_i * (p.tiles_i + 1) / 2 : p.tiles_i * p.tiles_j;
  const unsigned grid = (unsigned)((long)(p.nsplit + 7) / 8 * 8 * tiles);
  // algorithmic flops 2 M1 M2 K (the full product, whatever the triangle or split); the
  // f16 MFMA flops of the tiles computed go to "gemm_mixed_tn.mfma"
  const double frac = tri ? (double)tiles / ((double)p.tiles_i * p.tiles_j) : 1.0;
  timing_add_work("gemm_mixed_tn.mfma", frac * 2.0 * (nterms >= 4 ? 4 : 3) * (double)M1 * M2 * K);
  Timed t("gemm_mixed_tn", 2.0 * (double)M1 * M2 * K);
  if (nterms >= 4) hipLaunchKernelGGL(k_gemm_tnf<4>, dim3(grid), dim3(512), 0, s, p);
  else hipLaunchKernelGGL(k_gemm_tnf<3>, dim3(grid), dim3(512), 0, s, p);
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_tn_reduce, dim3(nblk((long)M1 * M2, 256)), dim3(256), 0, s, p.part, p.nsplit,
                     M1, M2, alpha, sa, sb, beta, C, ldc, tri);
  HIPCK(hipGetLastError());
}

void split_bands(const double *X, long ldx, long R, int M, const double *s, SplitBands &out,
                 hipStream_t st) {
  if (R <= 0 || M <= 0) throw Error(1, "split_bands: empty operand");
  const long Rp = (R + 255) / 256 * 256, KC = Rp / 16;
  const int Mp = (M + 255) / 256 * 256, Kpn = (M + 63) / 64 * 64;
  // the same buffers the products would split into (TN's first operand, NN's left factor)
  out.X = X; out.ldx = ldx; out.R = R; out.M = M; out.s = s;
  out.th = (_Float16 *)scratch("gemm.tnah", sizeof(_Float16) * (size_t)Mp * Rp);
  out.tl = (_Float16 *)scratch("gemm.tnal", sizeof(_Float16) * (size_t)Mp * Rp);
  out.nh = (_Float16 *)scratch("gemm.nnxh", sizeof(_Float16) * (size_t)Kpn * Rp);
  out.nl = (_Float16 *)scratch("gemm.nnxl", sizeof(_Float16) * (size_t)Kpn * Rp);
  Timed t("gemm_split", 8.0 * (double)R * M + 8.0 * (double)Rp * Mp);
  hipLaunchKernelGGL(k_split_both, dim3((unsigned)(Rp / 64), (unsigned)(Mp / 64)), dim3(256), 0, st, X,
                     ldx, R, M, s, KC, Kpn / 16, out.th, out.tl, out.nh, out.nl);
  HIPCK(hipGetLastError());
}

void maxscale(const unsigned long long *cmax, int cols, double *scale, hipStream_t s) {
  if (cols <= 0) return;
  hipLaunchKernelGGL(k_maxscale, dim3(nblk(cols, 256)), dim3(256), 0, s, cols, cmax, scale);
  HIPCK(hipGetLastError());
}

void normscale(const double *g, int cols, double dv, double *scale, hipStream_t s) {
  if (cols <= 0) return;
  hipLaunchKernelGGL(k_normscale, dim3(nblk(cols, 256)), dim3(256), 0, s, cols, g, dv, scale);
  HIPCK(hipGetLastError());
}

void gemm_mixed_tn(int M1, int M2, long K, double alpha, const double *A, long lda,
                   const double *sa, const double *B, long ldb, const double *sb, double beta,
                   double *C, long ldc, int nterms, hipStream_t s, const SplitBands *pa, int tri) {
  if (M1 <= 0 || M2 <= 0 || K <= 0) throw Error(1, "gemm_mixed_tn: empty product");
  if (pa || param("tn_presplit", 1.0) >= 1.0) {
    gemm_mixed_tn_presplit(M1, M2, K, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, nterms, s, pa,
                           tri);
    return;
  }
  TNArgs p{};
  p.M1 = M1; p.M2 = M2; p.K = K; p.A = A; p.B = B; p.lda = lda; p.ldb = ldb; p.sa = sa; p.sb = sb;
  p.tiles_i = (M1 + BM - 1) / BM;
  p.tiles_j = (M2 + BN - 1) / BN;
  const int tiles = p.tiles_i * p.tiles_j;
  // K per work-group: 4096 (FP32 accumulation depth), more if the FP32 partials would
  // exceed 2^29 values (2 GiB)
  const long max_split = std::max<long>(1, (1L << 29) / ((long)M1 * M2));
  long kper = std::max<long>(4096, (K + max_split - 1) / max_split);
  kper = (kper + BK - 1) / BK * BK;
  const int nsplit = (int)((K + kper - 1) / kper);
  p.kper = kper;
  p.nsplit = nsplit;
  p.nsplit_pad = (nsplit + 7) / 8 * 8;
  p.part = (float *)scratch("gemm.tnpart", sizeof(float) * (size_t)nsplit * M1 * M2);
  timing_add_work("gemm_mixed_tn.mfma", 2.0 * (nterms >= 4 ? 4 : 3) * (double)M1 * M2 * K);
  Timed t("gemm_mixed_tn", 2.0 * (double)M1 * M2 * K);  // algorithmic flops
  const unsigned grid = (unsigned)((long)p.nsplit_pad * tiles);
  if (nterms >= 4) hipLaunchKernelGGL(k_gemm_tn<4>, dim3(grid), dim3(512), 0, s, p);
  else hipLaunchKernelGGL(k_gemm_tn<3>, dim3(grid), dim3(512), 0, s, p);
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_tn_reduce, dim3(nblk((long)M1 * M2, 256)), dim3(256), 0, s, p.part, nsplit,
                     M1, M2, alpha, sa, sb, beta, C, ldc, 0);
  HIPCK(hipGetLastError());
}

void gemm_mixed_nn(long R, int K, int M2, double alpha, const double *X, long ldx,
                   const double *sx, const double *C, long ldc, double beta, double *Y, long ldy,
                   int nterms, hipStream_t s, const SplitBands *px, double *colsq, double dv,
                   bool c_upper) {
  if (colsq && beta != 0.0) throw Error(1, "gemm_mixed_nn: column norms need beta = 0");
  if (R <= 0 || M2 <= 0 || K <= 0) throw Error(1, "gemm_mixed_nn: empty product");
  const int Kp = (K + 63) / 64 * 64;  // whole 64-k split tiles (and 16-k stages)
  const int M2p = (M2 + NN_BN - 1) / NN_BN * NN_BN;
  const long Rp = (R + NN_BM - 1) / NN_BM * NN_BM;
  frag_rows_ok(Kp / 16, "gemm_mixed_nn");
  _Float16 *ch = (_Float16 *)scratch("gemm.nnch", sizeof(_Float16) * (size_t)Kp * M2p);
  _Float16 *cl = (_Float16 *)scratch("gemm.nncl", sizeof(_Float16) * (size_t)Kp * M2p);
  if (px && !(px->X == X && px->ldx == ldx && px->R == R && px->M == K && px->s == sx))
    throw Error(1, "gemm_mixed_nn: pre-split operand does not match X");
  _Float16 *xh = px ? px->nh : (_Float16 *)scratch("gemm.nnxh", sizeof(_Float16) * (size_t)Kp * Rp);
  _Float16 *xl = px ? px->nl : (_Float16 *)scratch("gemm.nnxl", sizeof(_Float16) * (size_t)Kp * Rp);
  double *tc = (double *)scratch("gemm.nntc", sizeof(double) * M2p);
  double *ys = (double *)scratch("gemm.nnys", sizeof(double) * M2p);
  hipLaunchKernelGGL(k_nn_cscale, dim3(M2), dim3(256), 0, s, C, ldc, K, M2, sx, alpha, tc, ys);
  hipLaunchKernelGGL(k_nn_prep, dim3(nblk((long)Kp * M2p, 256)), dim3(256), 0, s, C, ldc, K, M2, sx,
                     tc, Kp, M2p, ch, cl);
  HIPCK(hipGetLastError());
  if (!px) {
    Timed t("gemm_split", 8.0 * (double)R * K + 4.0 * (double)Rp * Kp);
    hipLaunchKernelGGL(k_split_t, dim3((unsigned)(Rp / 64), (unsigned)(Kp / 64)), dim3(256), 0, s, X,
                       ldx, R, K, sx, Kp, xh, xl);
    HIPCK(hipGetLastError());
  }
  NNArgs p{};
  p.R = R; p.M2 = M2; p.Kp = Kp; p.xh = xh; p.xl = xl; p.ch = ch; p.cl = cl;
  p.ys = ys; p.beta = beta; p.Y = Y; p.ldy = ldy;
  p.cupper = c_upper ? 1 : 0;
  // rows per work-group: 256 (one work-group per CU, the faster: 13.9 against 16.3 ms at
  // 2.1M x 1024 x 1024, beta = 0) or 128 (two per CU); Rp is a multiple of both
  const int rows = param("nn_rows", 256.0) >= 256.0 ? 256 : 128;
  p.row_blocks = (int)(Rp / rows);
  p.col_tiles = M2p / NN_BN;
  const long groups = (p.row_blocks + 7) / 8;
  const unsigned grid = (unsigned)(groups * 8 * p.col_tiles);
  double *part = nullptr;
  if (colsq) {
    part = (double *)scratch("gemm.nn.colsq", sizeof(double) * (size_t)p.row_blocks * M2);
    p.colsq = part;
  }
  double frac = 1.0;  // share of the K stages computed
  if (c_upper) {
    const long KC = Kp / 16;
    double done = 0.0;
    for (int ct = 0; ct < p.col_tiles; ct++) done += (double)std::min(KC, (long)(ct + 1) * NN_BN / 16);
    frac = done / ((double)KC * p.col_tiles);
  }
  {
    // algorithmic flops 2 R M2 K; the f16 MFMA flops of the stages computed beside them
    timing_add_work("gemm_mixed_nn.mfma", frac * 2.0 * (nterms >= 4 ? 4 : 3) * (double)R * M2 * K);
    Timed t("gemm_mixed_nn", 2.0 * (double)R * M2 * K);
    if (rows == 256) {
      if (nterms >= 4) hipLaunchKernelGGL((k_gemm_nn<4, 4, NN_NS>), dim3(grid), dim3(512), 0, s, p);
      else hipLaunchKernelGGL((k_gemm_nn<3, 4, NN_NS>), dim3(grid), dim3(512), 0, s, p);
    } else {
      if (nterms >= 4) hipLaunchKernelGGL((k_gemm_nn<4, 2, 3>), dim3(grid), dim3(256), 0, s, p);
      else hipLaunchKernelGGL((k_gemm_nn<3, 2, 3>), dim3(grid), dim3(256), 0, s, p);
    }
    HIPCK(hipGetLastError());
  }
  if (colsq) launch_reduce_partials(part, p.row_blocks, M2, dv, colsq, s);
}

}  // namespace kukan
