// batch.hip — batching pattern: many independent periodic systems of different grid sizes
// advanced together.  Per system and step: H psi for all bands, one preconditioned
// steepest-descent band update with a 2x2 Rayleigh-Ritz per band (the first, reset iteration
// of kscg_r, kscg8f.F90:98-472 and :1197-1230), the orthonormalisation of the bands (FP64
// Cholesky-QR, the Q of orthogonalization8f.F90:93-376), then the Hartree CG solve of
// rho = sum_l psi_l^2 (scf_hartree8f.F90).
//
// Batched mode: every stencil pass of all systems is one persistent launch (the items of all
// systems in one work list), per-band arithmetic runs as launches over a global band table,
// the Hartree CG of all systems advances as device-side iteration blocks with per-system stop
// flags, and the band step and the Hartree block are captured once as HIP graphs and
// replayed.  Loop mode runs the same kernels system by system (one launch per system and
// pass, no graphs): the comparison the batching is measured against.  Both modes compute
// bit-identical results.  The stencil tile (points per thread) and the persistent grid size
// are chosen by timing the candidates on first use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kukan_hip.h"
#include "kukan_internal.hpp"
#include "solvers.hpp"

namespace kukan {
namespace {

// One band of one system in the global band table.
struct BandRef {
  double *psi, *hpsi, *p, *hp;  // band columns
  const double *part;           // this band's stencil+projector partial sums
  long n;                       // grid points
  int nparts;
  double dv;
};
constexpr int BCH = 64;  // work-groups per band for band reductions

// which = 1: r0 = <psi|p>, r1 = <p|p>; which = 2: r0 = <psi|hp>  (times dV, fixed order)
__global__ __launch_bounds__(256) void k_band_dots(const BandRef *bands, int which, double *part) {
  __shared__ double red[2][4];
  const BandRef B = bands[blockIdx.y];
  const long chunk = (B.n + BCH - 1) / BCH;
  const long i0 = blockIdx.x * chunk, i1 = min(B.n, i0 + chunk);
  double s0 = 0.0, s1 = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double x = B.psi[i];
    if (which == 1) {
      const double q = B.p[i];
      s0 += x * q;
      s1 += q * q;
    } else {
      s0 += x * B.hp[i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *o = part + ((long)blockIdx.y * BCH + blockIdx.x) * 2;
    o[0] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    o[1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  }
}
__global__ void k_band_fold(const BandRef *bands, int nbt, const double *part, double *r0,
                            double *r1) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbt) return;
  double s0 = 0.0, s1 = 0.0;
  for (int c = 0; c < BCH; c++) {
    s0 += part[((long)b * BCH + c) * 2];
    s1 += part[((long)b * BCH + c) * 2 + 1];
  }
  r0[b] = bands[b].dv * s0;
  if (r1) r1[b] = bands[b].dv * s1;
}
// res[b] = dV * sum of the band's stencil/projector partials (<in|H in>)
__global__ void k_band_parts(const BandRef *bands, int nbt, double *res) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbt) return;
  const BandRef B = bands[b];
  double s = 0.0;
  for (int c = 0; c < B.nparts; c++) s += B.part[c];
  res[b] = B.dv * s;
}
// p <- (p - d psi) / sqrt(<p|p> - d^2)  (S = 1, psi normalised)
__global__ __launch_bounds__(256) void k_band_orth(const BandRef *bands, const double *d,
                                                   const double *pp) {
  const BandRef B = bands[blockIdx.y];
  const double db = d[blockIdx.y], n2 = pp[blockIdx.y] - db * db;
  const double f = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < B.n; i += stride) B.p[i] = f * (B.p[i] - db * B.psi[i]);
}
// 2x2 Rayleigh-Ritz of (psi, p) (kscg8f.F90:1197-1230, S = 1):
//   gamma = (eps + pap)/2 - sqrt(((eps - pap)/2)^2 + sap^2); lower eigenvector (a, c), a >= 0
__global__ void k_band_rr(int nbt, const double *eps, const double *pap, const double *sap,
                          double *ab, double *sval) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbt) return;
  const double e = eps[b], q = pap[b], s = sap[b];
  const double g = 0.5 * (e + q) - sqrt(0.25 * (e - q) * (e - q) + s * s);
  double a, c;
  if (fabs(e - g) >= fabs(q - g)) {
    a = -s;
    c = e - g;
  } else {
    a = q - g;
    c = -s;
  }
  double nn = sqrt(a * a + c * c);
  if (nn == 0.0) {
    a = 1.0;
    c = 0.0;
    nn = 1.0;
  }
  a /= nn;
  c /= nn;
  if (a < 0.0) {
    a = -a;
    c = -c;
  }
  ab[2 * b] = a;
  ab[2 * b + 1] = c;
  sval[b] = g;
}
// psi <- a psi + c p ; hpsi <- a hpsi + c hp
__global__ __launch_bounds__(256) void k_band_update(const BandRef *bands, const double *ab) {
  const BandRef B = bands[blockIdx.y];
  const double a = ab[2 * blockIdx.y], c = ab[2 * blockIdx.y + 1];
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < B.n; i += stride) {
    B.psi[i] = a * B.psi[i] + c * B.p[i];
    B.hpsi[i] = a * B.hpsi[i] + c * B.hp[i];
  }
}

// ------------------------------------------------------------------ batched Hartree CG
// state[0] = done, state[1] = iterations; sc = {r1r, pap, r2r, mean b}
struct HSys {
  double *rho, *vh, *b, *r, *p, *ap, *avhs, *sc, *part;
  int *state;
  long n;
  int np;
  double dvo;
};
constexpr int HCH = 64;

__global__ __launch_bounds__(256) void k_hb_sum(const HSys *hs, int which, double *part) {
  __shared__ double red[4];
  const HSys H = hs[blockIdx.y];
  const long chunk = (H.n + HCH - 1) / HCH;
  const long i0 = blockIdx.x * chunk, i1 = min(H.n, i0 + chunk);
  double s = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    if (which == 0) {
      const double bi = 2.0 * M_PI * H.rho[i];
      H.b[i] = bi;
      s += bi;
    } else {
      s += H.r[i] * H.r[i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(long)blockIdx.y * HCH + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}
__global__ void k_hb_fold(const HSys *hs, int ns, const double *part, int slot, int mean) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  double s = 0.0;
  for (int c = 0; c < HCH; c++) s += part[(long)i * HCH + c];
  hs[i].sc[slot] = mean ? s / (double)hs[i].n : s;
  if (slot == 0) {
    hs[i].state[0] = 0;
    hs[i].state[1] = 0;
  }
}
// b -= mean ; avhs = A v (in ap) ; r = b - avhs ; p = r
__global__ __launch_bounds__(256) void k_hb_init(const HSys *hs) {
  const HSys H = hs[blockIdx.y];
  const double m = H.sc[3];
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < H.n; i += stride) {
    const double bi = H.b[i] - m;
    H.b[i] = bi;
    H.avhs[i] = H.ap[i];
    const double ri = bi - H.ap[i];
    H.r[i] = ri;
    H.p[i] = ri;
  }
}
__global__ void k_hb_begin(const HSys *hs, int ns, int *alldone) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *alldone = 1;
  if (i >= ns) return;
  const HSys H = hs[i];
  if (H.state[0]) return;
  H.state[1] += 1;
  if (H.state[1] > 1) H.sc[0] = H.sc[2];
}
__global__ void k_hb_pap(const HSys *hs, int ns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const HSys H = hs[i];
  if (H.state[0]) return;
  double s = 0.0;
  for (int c = 0; c < H.np; c++) s += H.part[c];
  H.sc[1] = s;
}
__global__ __launch_bounds__(256) void k_hb_update(const HSys *hs, double *part) {
  __shared__ double red[4];
  const HSys H = hs[blockIdx.y];
  if (H.state[0]) return;
  const double alph = H.sc[0] / H.sc[1];
  const long chunk = (H.n + HCH - 1) / HCH;
  const long i0 = blockIdx.x * chunk, i1 = min(H.n, i0 + chunk);
  double s = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    H.vh[i] = H.vh[i] + alph * H.p[i];
    H.avhs[i] = H.avhs[i] + alph * H.ap[i];
    const double ri = H.b[i] - H.avhs[i];
    H.r[i] = ri;
    s += ri * ri;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(long)blockIdx.y * HCH + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}
__global__ void k_hb_r2r(const HSys *hs, int ns, const double *part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const HSys H = hs[i];
  if (H.state[0]) return;
  double s = 0.0;
  for (int c = 0; c < HCH; c++) s += part[(long)i * HCH + c];
  H.sc[2] = s;
}
__global__ __launch_bounds__(256) void k_hb_dir(const HSys *hs, int ncgres) {
  const HSys H = hs[blockIdx.y];
  if (H.state[0]) return;
  const double beta = (H.state[1] % ncgres == 0) ? 0.0 : H.sc[2] / H.sc[0];
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < H.n; i += stride) H.p[i] = H.r[i] + beta * H.p[i];
}
__global__ void k_hb_check(const HSys *hs, int ns, double eps, int ncgmin, int ncgmax,
                           int *alldone) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const HSys H = hs[i];
  if (!H.state[0]) {
    const int loop = H.state[1];
    const double dp = sqrt(H.sc[2] * H.dvo);
    if (loop >= ncgmin && (loop >= ncgmax || dp <= eps)) H.state[0] = 1;
  }
  if (!H.state[0]) atomicAnd(alldone, 0);
}

inline unsigned nb_(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ============================================================================= Batch
// ---------------------------------------------------------------- orthonormalisation
// One FP64 Cholesky-QR per system after the band update (S = 1): G = dV Psi^T Psi = R^T R,
// Psi <- Psi R^-1, the same Q as the reference's Gram-Schmidt (orthogonalization8f.F90:93-376)
// in exact arithmetic.  The work of all systems runs in one launch per kernel over unit
// tables (system-major), so loop mode launches the same kernels on one system's units.
constexpr int GT = 32;          // band tile of the Gram partials
constexpr long GCH = 16384;     // points per Gram work-group
constexpr int OMAX = 128;       // bands per system (the Cholesky runs in 128 KB of LDS)
struct OSys {
  double *psi;
  long n;
  int M, ntile, nchunk, ntp;
  double dv;
  double *G, *Rinv, *part;  // M x M, M x M, [tile pair][chunk][GT x GT]
  int *flag;                // 1: a pivot was not positive
};
struct GUnit {
  int sys, tp, ti, tj, chunk;
};
struct AUnit {
  int sys;
  long p0;  // first point of the 64-point tile
};

// Gram partial of band tiles (ti, tj) over one chunk of points: 4 outputs per thread
__global__ __launch_bounds__(256) void k_bgram_part(const OSys *os, const GUnit *units) {
  __shared__ double A[GT][65], Bt[GT][65];
  const GUnit u = units[blockIdx.x];
  const OSys S = os[u.sys];
  const long c0 = (long)u.chunk * GCH, c1 = min(S.n, c0 + GCH);
  const int t = threadIdx.x, i = t & 31, j0 = t >> 5;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long p0 = c0; p0 < c1; p0 += 64) {
    for (int e = t; e < GT * 64; e += 256) {
      const int b = e >> 6, q = e & 63;
      const long pt = p0 + q;
      const int ba = u.ti * GT + b, bb = u.tj * GT + b;
      A[b][q] = (ba < S.M && pt < c1) ? S.psi[(long)ba * S.n + pt] : 0.0;
      Bt[b][q] = (bb < S.M && pt < c1) ? S.psi[(long)bb * S.n + pt] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < 64; q++) {
      const double a = A[i][q];
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] += a * Bt[j0 + 8 * k][q];
    }
    __syncthreads();
  }
  double *o = S.part + ((long)u.tp * S.nchunk + u.chunk) * (GT * GT);
#pragma unroll
  for (int k = 0; k < 4; k++) o[i + GT * (j0 + 8 * k)] = acc[k];
}

// G tile (ti, tj) = dV sum over chunks in order
__global__ __launch_bounds__(256) void k_bgram_reduce(const OSys *os, const GUnit *units) {
  const GUnit u = units[blockIdx.x];  // chunk field unused: one unit per tile pair
  const OSys S = os[u.sys];
  for (int e = threadIdx.x; e < GT * GT; e += 256) {
    const int i = e % GT, j = e / GT, gi = u.ti * GT + i, gj = u.tj * GT + j;
    double s = 0.0;
    for (int c = 0; c < S.nchunk; c++) s += S.part[((long)u.tp * S.nchunk + c) * (GT * GT) + e];
    if (gi < S.M && gj < S.M) S.G[gi + (long)S.M * gj] = S.dv * s;
  }
}

// upper Cholesky factor of G in LDS, then its inverse (one column per thread), per system
__global__ __launch_bounds__(256) void k_bchol_inv(const OSys *os) {
  __shared__ double R[OMAX * OMAX];
  __shared__ int bad;
  const OSys S = os[blockIdx.x];
  const int M = S.M, t = threadIdx.x;
  for (int e = t; e < M * M; e += 256) R[e] = S.G[e];
  if (t == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < M; k++) {
    if (t == 0) {
      const double d = R[k + M * k];
      if (!(d > 0.0)) bad = 1;
      R[k + M * k] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double rkk = R[k + M * k];
    for (int j = k + 1 + t; j < M; j += 256) R[k + M * j] /= rkk;
    __syncthreads();
    const int m = M - k - 1;  // the trailing upper triangle, (i, j) with k < i <= j
    for (int e = t; e < m * m; e += 256) {
      const int i = k + 1 + e % m, j = k + 1 + e / m;
      if (i <= j) R[i + M * j] -= R[k + M * i] * R[k + M * j];
    }
    __syncthreads();
  }
  // X = R^-1 (upper): column j by back substitution, written to Rinv (zero below the diagonal)
  for (int j = t; j < M; j += 256) {
    double *X = S.Rinv + (long)M * j;
    for (int i = j + 1; i < M; i++) X[i] = 0.0;
    X[j] = 1.0 / R[j + M * j];
    for (int i = j - 1; i >= 0; i--) {
      double s = 0.0;
      for (int k = i + 1; k <= j; k++) s += R[i + M * k] * X[k];
      X[i] = -s / R[i + M * i];
    }
  }
  if (t == 0) *S.flag = bad;
}

// Psi <- Psi R^-1 on a tile of 64 points: the tile's bands staged in LDS, in place
__global__ __launch_bounds__(256) void k_bortho_apply(const OSys *os, const AUnit *units) {
  __shared__ double X[OMAX][64];
  const AUnit u = units[blockIdx.x];
  const OSys S = os[u.sys];
  const int t = threadIdx.x, q = t & 63;
  const long pt = u.p0 + q;
  for (int e = t; e < S.M * 64; e += 256) {
    const int b = e >> 6, qq = e & 63;
    const long p = u.p0 + qq;
    X[b][qq] = p < S.n ? S.psi[(long)b * S.n + p] : 0.0;
  }
  __syncthreads();
  if (pt >= S.n) return;
  for (int j = t >> 6; j < S.M; j += 4) {
    const double *rc = S.Rinv + (long)S.M * j;
    double y = 0.0;
    for (int i = 0; i <= j; i++) y += X[i][q] * rc[i];
    S.psi[(long)j * S.n + pt] = y;
  }
}

struct BatchSystem {
  int n = 0, M = 0, natom = 0, num_list = 0, band0 = 0;
  double h = 0;
  Grid g;
  std::unique_ptr<Ctx> ctx;  // projector plan and dij
  double *veff = nullptr, *vnl = nullptr, *dij = nullptr, *psi = nullptr, *hpsi = nullptr,
         *p = nullptr, *hp = nullptr, *rp = nullptr, *coef = nullptr, *part = nullptr,
         *ones = nullptr;
  int *lst = nullptr;
  double *rho = nullptr, *vh = nullptr, *hb = nullptr, *hr = nullptr, *hpv = nullptr,
         *hap = nullptr, *havhs = nullptr, *hsc = nullptr, *hpart = nullptr;
  int *hstate = nullptr;
  long nparts = 0, sparts = 0, hparts = 0;
  std::vector<void *> owned;
  template <class T>
  T *alloc(size_t count) {
    void *q = nullptr;
    HIPCK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    owned.push_back(q);
    return (T *)q;
  }
  ~BatchSystem() {
    for (void *q : owned) (void)hipFree(q);
  }
};

struct Batch {
  std::vector<std::unique_ptr<BatchSystem>> sys;
  int nbt = 0;
  BandRef *bands = nullptr, *pbands = nullptr;  // <psi|H psi> partials / <p|H p> partials
  double *bpart = nullptr, *eps = nullptr, *d = nullptr, *pp = nullptr, *sap = nullptr,
         *pap = nullptr, *ab = nullptr, *sval = nullptr, *hpart = nullptr;
  HSys *hs = nullptr;
  int *alldone = nullptr;
  std::vector<void *> owned;
  bool ready = false;
  int P = 2, grid = 1024;
  // plans over all systems [0] and per system [1 + i]; kinds: H psi, H p, precond, Hartree
  std::vector<StencilBatchPlan> plan[5];
  hipGraphExec_t g_band = nullptr, g_hart = nullptr;
  int hart_blk = 16;
  // orthonormalisation: per-system descriptors and unit tables (system-major) with each
  // system's first unit, so loop mode launches one system's range
  bool ortho = true;
  OSys *osys = nullptr;
  GUnit *gunits = nullptr, *runits = nullptr;
  AUnit *aunits = nullptr;
  std::vector<int> g0, r0, a0;  // ns + 1 offsets each
  int *oflag = nullptr;
  ~Batch() {
    if (g_band) (void)hipGraphExecDestroy(g_band);
    if (g_hart) (void)hipGraphExecDestroy(g_hart);
    for (void *q : owned) (void)hipFree(q);
  }
  template <class T>
  T *alloc(size_t count) {
    void *q = nullptr;
    HIPCK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    owned.push_back(q);
    return (T *)q;
  }
};

namespace {

std::vector<Batch *> &batches() {
  static std::vector<Batch *> b;
  return b;
}
Batch *get_batch(void *h) {
  for (Batch *b : batches())
    if (b == h) return b;
  throw Error(1, "batch: unknown handle");
}

StencilCall call_for(BatchSystem &S, int kind, const double *eps_band) {
  StencilCall c;
  c.g = &S.g;
  c.ld_in = c.ld_out = S.g.n();
  c.nb = S.M;
  if (kind == 0 || kind == 1) {  // H psi / H p with <in|H in> partials
    c.in = kind == 0 ? S.psi : S.p;
    c.out = kind == 0 ? S.hpsi : S.hp;
    c.veff = S.veff;
    c.dot_partial = S.part;
    c.dot_ld = S.nparts;
  } else if (kind == 2) {  // K r with r = -(H psi - eps psi)
    c.in = S.hpsi; c.in2 = S.psi; c.eps = eps_band; c.sign = -1.0; c.out = S.p;
  } else if (kind == 3) {  // Hartree: A p with <p|A p> partials, skipped once converged
    c.in = S.hpv; c.out = S.hap; c.nb = 1; c.nfh_coeffs = 4; c.dot_partial = S.hpart;
    c.dot_ld = S.hparts; c.skip = S.hstate;
  } else {  // Hartree start: A v
    c.in = S.vh; c.out = S.hap; c.nb = 1; c.nfh_coeffs = 4;
  }
  return c;
}

void build_plans(Batch &B) {
  const int ns = (int)B.sys.size();
  for (int kind = 0; kind < 5; kind++) {
    B.plan[kind].clear();
    std::vector<StencilCall> all;
    for (int i = 0; i < ns; i++) all.push_back(call_for(*B.sys[i], kind, B.eps + B.sys[i]->band0));
    const bool res = kind == 2;
    const std::string tag = "batch" + std::to_string((long)(intptr_t)&B) + ".k" + std::to_string(kind);
    B.plan[kind].push_back(plan_stencil_batch(all, res, B.P, tag + ".all"));
    for (int i = 0; i < ns; i++)
      B.plan[kind].push_back(plan_stencil_batch({all[i]}, res, B.P, tag + ".s" + std::to_string(i)));
  }
}

// ---------------------------------------------------------------- one band step
// systems [s0, s1); batched: one launch per pass over the whole range
void band_step(Batch &B, int s0, int s1, bool batched, hipStream_t s) {
  const int b0 = B.sys[s0]->band0;
  const int b1 = s1 < (int)B.sys.size() ? B.sys[s1]->band0 : B.nbt;
  const int nb = b1 - b0;
  auto pass = [&](int kind) {
    if (batched) {
      launch_stencil_batch(B.plan[kind][0], B.grid, s);
    } else {
      for (int i = s0; i < s1; i++) launch_stencil_batch(B.plan[kind][1 + i], B.grid, s);
    }
  };
  auto nonlocal = [&](bool on_p) {
    for (int i = s0; i < s1; i++) {
      BatchSystem &S = *B.sys[i];
      Ctx &c = *S.ctx;
      const long n1 = (long)c.pl.nprjmx * c.pl.num_ppcell;
      c.proj(on_p ? S.p : S.psi, S.g.n(), S.M, S.rp, n1);
      launch_coef(c.pl, S.rp, n1, S.M, c.dij, false, 1, 1.0 / S.g.dv(), S.coef, n1, s, true);
    }
  };
  auto expand = [&](bool on_p) {
    for (int i = s0; i < s1; i++) {
      BatchSystem &S = *B.sys[i];
      launch_expand_rows(S.ctx->pl, S.coef, S.M, on_p ? S.hp : S.hpsi, S.g.n(), on_p ? S.p : S.psi,
                         S.g.n(), S.part, S.nparts, (int)S.sparts, s);
    }
  };
  const dim3 gb(BCH, nb), gu(16, nb);
  // eps = dV <psi|H psi>
  nonlocal(false);
  pass(0);
  expand(false);
  hipLaunchKernelGGL(k_band_parts, dim3(nb_(nb, 64)), dim3(64), 0, s, B.bands + b0, nb, B.eps + b0);
  // p = K r, orthogonal to psi and normalised
  pass(2);
  hipLaunchKernelGGL(k_band_dots, gb, dim3(256), 0, s, B.bands + b0, 1, B.bpart + (long)b0 * BCH * 2);
  hipLaunchKernelGGL(k_band_fold, dim3(nb_(nb, 64)), dim3(64), 0, s, B.bands + b0, nb,
                     B.bpart + (long)b0 * BCH * 2, B.d + b0, B.pp + b0);
  hipLaunchKernelGGL(k_band_orth, gu, dim3(256), 0, s, B.bands + b0, B.d + b0, B.pp + b0);
  // H p, pap = dV <p|H p>, sap = dV <psi|H p>
  nonlocal(true);
  pass(1);
  expand(true);
  hipLaunchKernelGGL(k_band_parts, dim3(nb_(nb, 64)), dim3(64), 0, s, B.bands + b0, nb, B.pap + b0);
  hipLaunchKernelGGL(k_band_dots, gb, dim3(256), 0, s, B.bands + b0, 2, B.bpart + (long)b0 * BCH * 2);
  hipLaunchKernelGGL(k_band_fold, dim3(nb_(nb, 64)), dim3(64), 0, s, B.bands + b0, nb,
                     B.bpart + (long)b0 * BCH * 2, B.sap + b0, nullptr);
  // 2x2 Rayleigh-Ritz and the band update
  hipLaunchKernelGGL(k_band_rr, dim3(nb_(nb, 64)), dim3(64), 0, s, nb, B.eps + b0, B.pap + b0,
                     B.sap + b0, B.ab + 2L * b0, B.sval + b0);
  hipLaunchKernelGGL(k_band_update, gu, dim3(256), 0, s, B.bands + b0, B.ab + 2L * b0);
  // Cholesky-QR of each system's bands
  if (B.ortho) {
    const int ng = B.g0[s1] - B.g0[s0], nr = B.r0[s1] - B.r0[s0], na = B.a0[s1] - B.a0[s0];
    if (ng > 0) hipLaunchKernelGGL(k_bgram_part, dim3(ng), dim3(256), 0, s, B.osys, B.gunits + B.g0[s0]);
    if (nr > 0) hipLaunchKernelGGL(k_bgram_reduce, dim3(nr), dim3(256), 0, s, B.osys, B.runits + B.r0[s0]);
    hipLaunchKernelGGL(k_bchol_inv, dim3(s1 - s0), dim3(256), 0, s, B.osys + s0);
    if (na > 0) hipLaunchKernelGGL(k_bortho_apply, dim3(na), dim3(256), 0, s, B.osys, B.aunits + B.a0[s0]);
  }
  HIPCK(hipGetLastError());
}

// ---------------------------------------------------------------- Hartree
void hartree_start(Batch &B, int s0, int s1, bool batched, hipStream_t s) {
  const int ns = s1 - s0;
  for (int i = s0; i < s1; i++) {
    BatchSystem &S = *B.sys[i];
    launch_fill(S.g.n(), 0.0, S.rho, s);
    charge(S.g.n(), S.M, S.psi, S.ones, S.rho);
  }
  HSys *hs = B.hs + s0;
  double *part = B.hpart + (long)s0 * HCH;
  hipLaunchKernelGGL(k_hb_sum, dim3(HCH, ns), dim3(256), 0, s, hs, 0, part);
  hipLaunchKernelGGL(k_hb_fold, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns, part, 3, 1);
  if (batched) {
    launch_stencil_batch(B.plan[4][0], B.grid, s);
  } else {
    for (int i = s0; i < s1; i++) launch_stencil_batch(B.plan[4][1 + i], B.grid, s);
  }
  hipLaunchKernelGGL(k_hb_init, dim3(16, ns), dim3(256), 0, s, hs);
  hipLaunchKernelGGL(k_hb_sum, dim3(HCH, ns), dim3(256), 0, s, hs, 1, part);
  hipLaunchKernelGGL(k_hb_fold, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns, part, 0, 0);
  HIPCK(hipGetLastError());
}

void hartree_iter(Batch &B, int s0, int s1, bool batched, int ncgmin, int ncgmax, double eps,
                  hipStream_t s) {
  constexpr int ncgres = 801;  // restart period of parameters.inp
  const int ns = s1 - s0;
  HSys *hs = B.hs + s0;
  double *part = B.hpart + (long)s0 * HCH;
  hipLaunchKernelGGL(k_hb_begin, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns, B.alldone);
  if (batched) {
    launch_stencil_batch(B.plan[3][0], B.grid, s);
  } else {
    for (int i = s0; i < s1; i++) launch_stencil_batch(B.plan[3][1 + i], B.grid, s);
  }
  hipLaunchKernelGGL(k_hb_pap, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns);
  hipLaunchKernelGGL(k_hb_update, dim3(HCH, ns), dim3(256), 0, s, hs, part);
  hipLaunchKernelGGL(k_hb_r2r, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns, part);
  hipLaunchKernelGGL(k_hb_dir, dim3(16, ns), dim3(256), 0, s, hs, ncgres);
  hipLaunchKernelGGL(k_hb_check, dim3(nb_(ns, 64)), dim3(64), 0, s, hs, ns, eps, ncgmin, ncgmax,
                     B.alldone);
  HIPCK(hipGetLastError());
}

int alldone(Batch &B, hipStream_t s) {
  int h = 0;
  HIPCK(hipMemcpyAsync(&h, B.alldone, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  return h;
}

void hartree_solve(Batch &B, int s0, int s1, bool batched, int ncgmax, double eps,
                   hipStream_t s) {
  hartree_start(B, s0, s1, batched, s);
  const int blk = B.hart_blk;
  for (int done = 0, it = 0; !done && it < ncgmax; it += blk) {
    if (batched) {
      HIPCK(hipGraphLaunch(B.g_hart, s));
    } else {
      for (int k = 0; k < blk; k++) hartree_iter(B, s0, s1, false, 1, ncgmax, eps, s);
    }
    done = alldone(B, s);
  }
}

template <class F>
hipGraphExec_t capture(hipStream_t s, F &&f) {
  hipGraph_t graph;
  lib().capturing = true;
  try {
    HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    f();
    HIPCK(hipStreamEndCapture(s, &graph));
  } catch (...) {
    lib().capturing = false;
    hipGraph_t g2;
    (void)hipStreamEndCapture(s, &g2);
    throw;
  }
  lib().capturing = false;
  hipGraphExec_t exec;
  HIPCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HIPCK(hipGraphDestroy(graph));
  return exec;
}

// Tile shape and persistent grid: time the batched H psi pass for each candidate.
void autotune(Batch &B, hipStream_t s) {
  const int Ps[3] = {1, 2, 4};
  const int grids[4] = {512, 1024, 2048, 4096};
  float best = 1e30f;
  int bp = 2, bg = 1024;
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  for (int P : Ps) {
    bool fits = true;
    for (auto &sp : B.sys) fits = fits && (long)sp->g.nx * std::max(1, std::min(sp->g.ny, 512 * P / sp->g.nx)) <= 512L * P;
    if (!fits) continue;
    B.P = P;
    std::vector<StencilCall> all;
    for (auto &sp : B.sys) all.push_back(call_for(*sp, 0, nullptr));
    StencilBatchPlan plan = plan_stencil_batch(all, false, P, "batch.tune");
    for (int G : grids) {
      launch_stencil_batch(plan, G, s);  // warm
      HIPCK(hipEventRecord(e0, s));
      for (int r = 0; r < 3; r++) launch_stencil_batch(plan, G, s);
      HIPCK(hipEventRecord(e1, s));
      HIPCK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bp = P;
        bg = G;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  B.P = bp;
  B.grid = bg;
  lib().params["batch_tile_p"] = bp;
  lib().params["batch_grid"] = bg;
  lib().params["batch_tune_ms"] = best / 3.0;
}

void prepare(Batch &B) {
  if (B.ready) return;
  hipStream_t s = lib().stream;
  const int ns = (int)B.sys.size();
  if (ns == 0) throw Error(1, "batch: no systems");
  B.nbt = 0;
  for (auto &sp : B.sys) {
    sp->band0 = B.nbt;
    B.nbt += sp->M;
  }
  const int nbt = B.nbt;
  B.bands = B.alloc<BandRef>(nbt);
  B.bpart = B.alloc<double>((size_t)nbt * BCH * 2);
  B.eps = B.alloc<double>(nbt);
  B.d = B.alloc<double>(nbt);
  B.pp = B.alloc<double>(nbt);
  B.sap = B.alloc<double>(nbt);
  B.pap = B.alloc<double>(nbt);
  B.ab = B.alloc<double>(2L * nbt);
  B.sval = B.alloc<double>(nbt);
  B.hs = B.alloc<HSys>(ns);
  B.hpart = B.alloc<double>((size_t)ns * HCH);
  B.alldone = B.alloc<int>(1);
  autotune(B, s);
  std::vector<BandRef> br(nbt);
  std::vector<HSys> hs(ns);
  for (int i = 0; i < ns; i++) {
    BatchSystem &S = *B.sys[i];
    const long n = S.g.n();
    S.sparts = stencil_batch_parts(S.g, 4, S.M, B.P);
    S.nparts = S.sparts + expand_nparts(S.ctx->pl);
    S.part = S.alloc<double>((size_t)S.nparts * S.M);
    S.hparts = stencil_batch_parts(S.g, 4, 1, B.P);
    S.hpart = S.alloc<double>(S.hparts);
    for (int l = 0; l < S.M; l++) {
      BandRef &r = br[S.band0 + l];
      r.psi = S.psi + n * l; r.hpsi = S.hpsi + n * l; r.p = S.p + n * l; r.hp = S.hp + n * l;
      r.part = S.part + (long)S.nparts * l; r.n = n; r.nparts = (int)S.nparts; r.dv = S.g.dv();
    }
    HSys &H = hs[i];
    H.rho = S.rho; H.vh = S.vh; H.b = S.hb; H.r = S.hr; H.p = S.hpv; H.ap = S.hap; H.avhs = S.havhs;
    H.sc = S.hsc; H.part = S.hpart; H.state = S.hstate; H.n = n; H.np = (int)S.hparts;
    H.dvo = 1.0 / (double)n;  // dV / Omega of the periodic cell
  }
  HIPCK(hipMemcpy(B.bands, br.data(), sizeof(BandRef) * nbt, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(B.hs, hs.data(), sizeof(HSys) * ns, hipMemcpyHostToDevice));
  // orthonormalisation tables (param "batch_ortho", default 1, read once here)
  B.ortho = param("batch_ortho", 1.0) >= 1.0;
  if (B.ortho) {
    std::vector<OSys> os(ns);
    std::vector<GUnit> gu, ru;
    std::vector<AUnit> au;
    B.oflag = B.alloc<int>(ns);
    B.g0.assign(1, 0);
    B.r0.assign(1, 0);
    B.a0.assign(1, 0);
    for (int i = 0; i < ns; i++) {
      BatchSystem &S = *B.sys[i];
      if (S.M > OMAX) throw Error(1, "batch: at most 128 bands per system");
      OSys &o = os[i];
      o.psi = S.psi;
      o.n = S.g.n();
      o.M = S.M;
      o.ntile = (S.M + GT - 1) / GT;
      o.nchunk = (int)((o.n + GCH - 1) / GCH);
      o.ntp = o.ntile * (o.ntile + 1) / 2;
      o.dv = S.g.dv();
      o.G = S.alloc<double>((size_t)S.M * S.M);
      o.Rinv = S.alloc<double>((size_t)S.M * S.M);
      o.part = S.alloc<double>((size_t)o.ntp * o.nchunk * GT * GT);
      o.flag = B.oflag + i;
      int tp = 0;
      for (int tj = 0; tj < o.ntile; tj++)
        for (int ti = 0; ti <= tj; ti++, tp++) {
          for (int c = 0; c < o.nchunk; c++) gu.push_back({i, tp, ti, tj, c});
          ru.push_back({i, tp, ti, tj, 0});
        }
      for (long p0 = 0; p0 < o.n; p0 += 64) au.push_back({i, p0});
      B.g0.push_back((int)gu.size());
      B.r0.push_back((int)ru.size());
      B.a0.push_back((int)au.size());
    }
    B.osys = B.alloc<OSys>(ns);
    B.gunits = B.alloc<GUnit>(gu.size());
    B.runits = B.alloc<GUnit>(ru.size());
    B.aunits = B.alloc<AUnit>(au.size());
    HIPCK(hipMemcpy(B.osys, os.data(), sizeof(OSys) * ns, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(B.gunits, gu.data(), sizeof(GUnit) * gu.size(), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(B.runits, ru.data(), sizeof(GUnit) * ru.size(), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(B.aunits, au.data(), sizeof(AUnit) * au.size(), hipMemcpyHostToDevice));
  }
  build_plans(B);
  B.ready = true;
}

}  // namespace

// ============================================================================= C-ABI helpers
void *batch_create() {
  Batch *b = new Batch();
  batches().push_back(b);
  return b;
}

void batch_destroy(void *h) {
  Batch *b = get_batch(h);
  batches().erase(std::find(batches().begin(), batches().end(), b));
  sync();
  delete b;
}

int batch_add(void *h, int n, double hh, int nbands, int natom, int num_list, const int *lstvec2,
              const int *natinf, const double *vnlocp, const double *dij, const double *veff,
              const double *psi) {
  Batch *B = get_batch(h);
  if (B->ready) throw Error(1, "batch: systems can only be added before the first step");
  if (n < 8 || n > 256) throw Error(1, "batch: grid edge must be in 8..256");
  if (nbands < 1 || natom < 1 || num_list < 1) throw Error(1, "batch: empty system");
  hipStream_t s = lib().stream;
  auto S = std::make_unique<BatchSystem>();
  S->n = n; S->h = hh; S->M = nbands; S->natom = natom; S->num_list = num_list;
  S->g = Grid{n, n, n, 4, 3, hh, hh, hh};
  const long N = S->g.n();
  S->veff = S->alloc<double>(N);
  S->vnl = S->alloc<double>((size_t)num_list * 9 * natom);
  S->dij = S->alloc<double>((size_t)81 * natom);
  S->lst = S->alloc<int>((size_t)num_list * natom);
  S->psi = S->alloc<double>((size_t)N * nbands);
  S->hpsi = S->alloc<double>((size_t)N * nbands);
  S->p = S->alloc<double>((size_t)N * nbands);
  S->hp = S->alloc<double>((size_t)N * nbands);
  S->rp = S->alloc<double>((size_t)9 * natom * nbands);
  S->coef = S->alloc<double>((size_t)9 * natom * nbands);
  S->ones = S->alloc<double>(nbands);
  S->rho = S->alloc<double>(N);
  S->vh = S->alloc<double>(N);
  S->hb = S->alloc<double>(N);
  S->hr = S->alloc<double>(N);
  S->hpv = S->alloc<double>(N);
  S->hap = S->alloc<double>(N);
  S->havhs = S->alloc<double>(N);
  S->hsc = S->alloc<double>(8);
  S->hstate = S->alloc<int>(2);
  auto up = [&](void *dst, const void *src, size_t bytes) {
    Arr a(src, bytes, Arr::In);
    HIPCK(hipMemcpyAsync(dst, a.v(), bytes, hipMemcpyDeviceToDevice, s));
    HIPCK(hipStreamSynchronize(s));
  };
  up(S->veff, veff, sizeof(double) * N);
  up(S->vnl, vnlocp, sizeof(double) * num_list * 9 * natom);
  up(S->dij, dij, sizeof(double) * 81 * natom);
  up(S->lst, lstvec2, sizeof(int) * num_list * natom);
  up(S->psi, psi, sizeof(double) * N * nbands);
  launch_fill(nbands, 1.0, S->ones, s);
  launch_fill(N, 0.0, S->vh, s);
  // projector plan: one species, 9 projectors, every atom on this rank
  std::vector<int> nprj(1, 9), indspe(natom, 1), natpri(natom, 2), naps(natom);  // natpri: "in"
  std::vector<int> nat = host_ints(natinf, natom);
  for (int a = 0; a < natom; a++) naps[a] = a + 1;
  S->ctx = std::make_unique<Ctx>(S->g);
  build_plan(S->ctx->pl, S->g, natom, 9, 1, natom, num_list, 1, 2, 1, nprj.data(), indspe.data(),
             nat.data(), natpri.data(), naps.data(), S->vnl, S->lst,
             "batch" + std::to_string((long)(intptr_t)B) + ".p" + std::to_string(B->sys.size()));
  S->ctx->dij = S->dij;
  S->ctx->q_zero = true;
  B->sys.push_back(std::move(S));
  sync();
  return (int)B->sys.size() - 1;
}

void batch_step(void *h, int batched, int ncgmax, double epsvh) {
  Batch *B = get_batch(h);
  prepare(*B);
  hipStream_t s = lib().stream;
  const int ns = (int)B->sys.size();
  if (batched) {
    Timed t("batch_step", 0.0);
    if (!B->g_band) B->g_band = capture(s, [&] { band_step(*B, 0, ns, true, s); });
    HIPCK(hipGraphLaunch(B->g_band, s));
    if (!B->g_hart)
      B->g_hart = capture(s, [&] {
        for (int k = 0; k < B->hart_blk; k++) hartree_iter(*B, 0, ns, true, 1, ncgmax, epsvh, s);
      });
    hartree_solve(*B, 0, ns, true, ncgmax, epsvh, s);
  } else {
    Timed t("batch_loop", 0.0);
    for (int i = 0; i < ns; i++) {
      band_step(*B, i, i + 1, false, s);
      hartree_solve(*B, i, i + 1, false, ncgmax, epsvh, s);
    }
  }
  sync();
  if (B->ortho) {
    std::vector<int> f(ns);
    HIPCK(hipMemcpy(f.data(), B->oflag, sizeof(int) * ns, hipMemcpyDeviceToHost));
    for (int i = 0; i < ns; i++)
      if (f[i]) throw Error(2, "batch: orthonormalisation of system " + std::to_string(i) + ": wf has negative norm");
  }
}

void batch_get(void *h, int i, double *psi, double *sval, double *vh, int *iters) {
  Batch *B = get_batch(h);
  if (i < 0 || i >= (int)B->sys.size()) throw Error(1, "batch: system index out of range");
  prepare(*B);
  BatchSystem &S = *B->sys[i];
  const long N = S.g.n();
  hipStream_t s = lib().stream;
  if (psi) {
    Arr a(psi, sizeof(double) * N * S.M, Arr::Out);
    HIPCK(hipMemcpyAsync(a.v(), S.psi, sizeof(double) * N * S.M, hipMemcpyDeviceToDevice, s));
    a.finish();
  }
  if (sval) {
    Arr a(sval, sizeof(double) * S.M, Arr::Out);
    HIPCK(hipMemcpyAsync(a.v(), B->sval + S.band0, sizeof(double) * S.M, hipMemcpyDeviceToDevice, s));
    a.finish();
  }
  if (vh) {
    Arr a(vh, sizeof(double) * N, Arr::Out);
    HIPCK(hipMemcpyAsync(a.v(), S.vh, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    a.finish();
  }
  if (iters) {
    int st[2];
    HIPCK(hipMemcpyAsync(st, S.hstate, sizeof(st), hipMemcpyDeviceToHost, s));
    sync();
    *iters = st[1];
  }
}

}  // namespace kukan
