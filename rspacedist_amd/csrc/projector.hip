// projector.hip — the projector tables and list vectors on the GPU (declared in
// include/kukan_hip.h): the double-grid weights (pseudocalc_01a), the atom information and the
// sphere list vectors of every atom (pseudocalc_01b), and the double-grid-filtered nonlocal
// projector table vnlocp (pseudocalc_07), for one rank.
//
// List vectors are integer results and match the reference bit for bit: the sphere tests and
// coordinates are evaluated without contraction to FMA, in the reference's operand order, and
// each atom's list is compacted by one work-group scanning its box in the reference's z-y-x
// order (ballot prefix sums), so the order of the entries is the reference's.  The
// neighbourhood test of a coarse point (does any of its (2 nfdg + 1)^3 neighbours lie inside
// the sphere?) takes the smallest squared offset along each axis separately: the rounded sum
// x^2 + y^2 + z^2 and its square root are non-decreasing in each term, so the smallest
// rounded distance over the neighbourhood is the one of the per-axis minima, and the test
// gives the reference's answer with 3 (2 nfdg + 1) instead of (2 nfdg + 1)^3 evaluations.
//
// vnlocp: the fine-grid projectors (one thread per fine point, all radial channels of the
// species from one sin/cos per Bessel term), then the separable double-grid filter as three
// passes (z, y, and x fused with the read-out at the list points), each output summed over its
// taps in the reference's order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "abi_common.hpp"
#include "kukan_hip.h"

using namespace kukan;
using namespace kukan::abi;

namespace {

constexpr int TB = 256;  // threads per work-group of the list scans (four waves)

__host__ __device__ inline int wrap1(int i, int n) { return ((i - 1) % n + n) % n + 1; }

struct Wrap {
  int nperi, nx, ny, nz;  // periodic extents (2 nxmax [* nmesh] ...)
  __device__ void operator()(int &i, int &j, int &k) const {
    if (nperi >= 1) {
      i = wrap1(i, nx);
      if (nperi >= 2) {
        j = wrap1(j, ny);
        if (nperi == 3) k = wrap1(k, nz);
      }
    }
  }
};

__device__ inline bool inside(int i, int j, int k, int nx, int ny, int nz) {
  return i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
}

// ============================================================================ pseudocalc_01a
// pseudocalc8f.F90:343-629: the Lagrange basis of node 0 among the 2 nfdg nodes around r
__global__ void k_dg_weights(int nfdg, int npmesh, double *w) {
#pragma clang fp contract(off)
  const int lo = -nfdg * npmesh, je = nfdg * npmesh - 1;
  const bool odd = npmesh % 2 == 1;
  const int js = odd ? lo + 1 : lo, ntmp1 = odd ? 0 : 1;
  const double tmp1 = odd ? 0.0 : 0.5;
  const double tmp = 1.0 / npmesh;
  const int j = js + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (odd && blockIdx.x == 0 && threadIdx.x == 0) w[0] = 0.0;
  if (j > je) return;
  const double r = fabs(j * tmp + tmp1 * tmp);
  int s = 0;
  for (int k = 1; k <= nfdg - 1; k++)
    if (abs(2 * j + ntmp1) > 2 * k * npmesh) s = k;
  double prod = 0.0, den = 1.0;
  bool first = true;
  for (int m = s + nfdg; m >= s - nfdg + 1; m--) {
    if (m == 0) continue;
    const double f = r - (double)m;
    prod = first ? f : prod * f;
    first = false;
    den *= (double)(-m);
  }
  w[j - lo] = prod / den;
}

// ============================================================================ pseudocalc_01b
struct ListArgs {
  int natom, nmesh, nfdg, nperi, nradmx;
  int nxmax, nymax, nzmax, ncpx, ncpy, ncpz, npxmax, npymax, npzmax;
  int num_list, num_list_d, num_atcell, num_ppcell, num_ppcell_d;
  int key_in, key_inps, key_out;
  double psctoff, xmax, ymax, zmax, dx, dy, dz;
  const int *indspe, *nradct;
  const double *radial, *atx, *aty, *atz;
  int *natx, *naty, *natz, *ndatx, *ndaty, *ndatz;
  int *natpri, *natprid, *natpri_inf, *naps, *napsd, *natinf, *natinfd, *natinfd_vloc;
  int *lstvec2, *lstx, *lsty, *lstz, *lstvecd2, *lstdx, *lstdy, *lstdz;
  int *flag;  // natom: an out atom's projector sphere reaches the grid
  int *err;   // 5 flags: num_atcell, num_ppcell, num_ppcell_d, num_list, num_list_d exceeded
};

__device__ inline double rc_of(const ListArgs &a, int na) {
  const int sp = a.indspe[na] - 1;
  return a.radial[(long)(a.nradct[sp] - 1) + (long)a.nradmx * sp];
}

// pseudocalc8f.F90:683-709: nearest coarse and dense points, in/out by the point's rank
__global__ void k_atom_info(ListArgs a) {
#pragma clang fp contract(off)
  const int na = blockIdx.x * blockDim.x + threadIdx.x;
  if (na >= a.natom) return;
  const double ddx = a.dx / a.nmesh, ddy = a.dy / a.nmesh, ddz = a.dz / a.nmesh;
  const int nx = (int)((a.atx[na] + a.xmax + 0.5 * a.dx) / a.dx) - a.nxmax;
  const int ny = (int)((a.aty[na] + a.ymax + 0.5 * a.dy) / a.dy) - a.nymax;
  const int nz = (int)((a.atz[na] + a.zmax + 0.5 * a.dz) / a.dz) - a.nzmax;
  a.natx[na] = nx;
  a.naty[na] = ny;
  a.natz[na] = nz;
  a.ndatx[na] = (int)((a.atx[na] + a.xmax + 0.5 * ddx) / ddx) - a.nxmax * a.nmesh;
  a.ndaty[na] = (int)((a.aty[na] + a.ymax + 0.5 * ddy) / ddy) - a.nymax * a.nmesh;
  a.ndatz[na] = (int)((a.atz[na] + a.zmax + 0.5 * ddz) / ddz) - a.nzmax * a.nmesh;
  // Fortran integer division truncates toward zero, as C++'s does
  const int i = (nx + a.nxmax - 1) / a.ncpx, j = (ny + a.nymax - 1) / a.ncpy, k = (nz + a.nzmax - 1) / a.ncpz;
  a.natpri[na] = (i == 0 && j == 0 && k == 0) ? a.key_in : a.key_out;
}

// the atom's offset from its coarse point (pseudocalc8f.F90:798-800)
struct CoarseBox {
  double ax, ay, az, thr;
  int nx, ny, nz;  // natx..
};
__device__ inline CoarseBox coarse_box(const ListArgs &a, int na) {
#pragma clang fp contract(off)
  CoarseBox b;
  b.nx = a.natx[na];
  b.ny = a.naty[na];
  b.nz = a.natz[na];
  b.ax = a.atx[na] - (b.nx * a.dx - 0.5 * a.dx);
  b.ay = a.aty[na] - (b.ny * a.dy - 0.5 * a.dy);
  b.az = a.atz[na] - (b.nz * a.dz - 0.5 * a.dz);
  b.thr = rc_of(a, na) * a.psctoff;
  return b;
}

// smallest rounded square offset (i + j) h - c over j = -nfdg..nfdg
__device__ inline double min_sq(int i, int nfdg, double h, double c) {
#pragma clang fp contract(off)
  double m = 0.0;
  for (int j = -nfdg; j <= nfdg; j++) {
    const double x = (i + j) * h - c;
    const double x2 = x * x;
    m = (j == -nfdg || x2 < m) ? x2 : m;
  }
  return m;
}

// coarse box point p (x fastest) of an atom: its neighbourhood meets the sphere, and the
// wrapped grid index it maps to
__device__ inline bool coarse_point(const ListArgs &a, const CoarseBox &b, long p, int &ix, int &iy, int &iz,
                                    int &i, int &j, int &k) {
#pragma clang fp contract(off)
  const int bx = 2 * a.npxmax, by = 2 * a.npymax;
  ix = (int)(p % bx) - a.npxmax + 1;
  iy = (int)((p / bx) % by) - a.npymax + 1;
  iz = (int)(p / ((long)bx * by)) - a.npzmax + 1;
  const double mx = min_sq(ix, a.nfdg, a.dx, b.ax), my = min_sq(iy, a.nfdg, a.dy, b.ay),
               mz = min_sq(iz, a.nfdg, a.dz, b.az);
  if (!(sqrt(mx + my + mz) < b.thr)) return false;
  i = ix + b.nx + a.nxmax;
  j = iy + b.ny + a.nymax;
  k = iz + b.nz + a.nzmax;
  Wrap{a.nperi, 2 * a.nxmax, 2 * a.nymax, 2 * a.nzmax}(i, j, k);
  return inside(i, j, k, a.ncpx, a.ncpy, a.ncpz);
}

// dense box point p: 0 outside the sphere, 1 inside radial(nradct), 2 in the shell up to
// radial(nradct) * psctoff; and whether its wrapped dense index lies in the dense grid
struct DenseBox {
  double bx, by, bz, rc, thr;
  int nx, ny, nz;
};
__device__ inline DenseBox dense_box(const ListArgs &a, int na) {
#pragma clang fp contract(off)
  DenseBox b;
  const double ddx = a.dx / a.nmesh, ddy = a.dy / a.nmesh, ddz = a.dz / a.nmesh;
  b.nx = a.ndatx[na];
  b.ny = a.ndaty[na];
  b.nz = a.ndatz[na];
  b.bx = a.atx[na] - (b.nx * ddx - 0.5 * ddx);
  b.by = a.aty[na] - (b.ny * ddy - 0.5 * ddy);
  b.bz = a.atz[na] - (b.nz * ddz - 0.5 * ddz);
  b.rc = rc_of(a, na);
  b.thr = b.rc * a.psctoff;
  return b;
}
__device__ inline int dense_point(const ListArgs &a, const DenseBox &b, long p, int &ix, int &iy, int &iz, int &i,
                                  int &j, int &k) {
#pragma clang fp contract(off)
  const int m = a.nmesh;
  const int bx = 2 * a.npxmax * m, by = 2 * a.npymax * m;
  ix = (int)(p % bx) - a.npxmax * m + 1;
  iy = (int)((p / bx) % by) - a.npymax * m + 1;
  iz = (int)(p / ((long)bx * by)) - a.npzmax * m + 1;
  const double ddx = a.dx / m, ddy = a.dy / m, ddz = a.dz / m;
  const double x = ix * ddx - b.bx, y = iy * ddy - b.by, z = iz * ddz - b.bz;
  const double r = sqrt(x * x + y * y + z * z);
  const int cls = r < b.rc ? 1 : (r < b.thr ? 2 : 0);
  if (!cls) return 0;
  i = ix + b.nx + a.nxmax * m;
  j = iy + b.ny + a.nymax * m;
  k = iz + b.nz + a.nzmax * m;
  Wrap{a.nperi, 2 * a.nxmax * m, 2 * a.nymax * m, 2 * a.nzmax * m}(i, j, k);
  return inside(i, j, k, a.ncpx * m, a.ncpy * m, a.ncpz * m) ? cls : 0;
}

// pseudocalc8f.F90:789-933: an out atom whose projector sphere reaches a grid point becomes
// inps; an atom whose dense sphere reaches a dense point has natprid = inps.  One work-group
// per atom.
__global__ void __launch_bounds__(TB) k_sphere_flags(ListArgs a) {
  const int na = blockIdx.x;
  int hit = 0;
  if (a.natpri[na] == a.key_out) {
    const CoarseBox b = coarse_box(a, na);
    const long np = 8L * a.npxmax * a.npymax * a.npzmax;
    for (long p = threadIdx.x; p < np && !hit; p += TB) {
      int ix, iy, iz, i, j, k;
      if (coarse_point(a, b, p, ix, iy, iz, i, j, k)) hit = 1;
    }
  }
  hit = __syncthreads_or(hit);
  const DenseBox db = dense_box(a, na);
  const long npd = 8L * a.npxmax * a.npymax * a.npzmax * a.nmesh * a.nmesh * a.nmesh;
  int hitd = 0;
  for (long p = threadIdx.x; p < npd && !hitd; p += TB) {
    int ix, iy, iz, i, j, k;
    if (dense_point(a, db, p, ix, iy, iz, i, j, k)) hitd = 1;
  }
  hitd = __syncthreads_or(hitd);
  if (threadIdx.x == 0) {
    a.flag[na] = hit;
    a.natprid[na] = hitd ? a.key_inps : a.key_out;
  }
}

// the running counters over atoms (pseudocalc8f.F90:716-726, 885-895, 937-948)
__global__ void k_counters(ListArgs a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int l = 0, iaps = 0, iapsd = 0;
  for (int na = 0; na < a.natom; na++) {
    a.natpri_inf[na] = 0;
    if (a.natpri[na] == a.key_in) {
      a.natpri_inf[na] = ++l;
      if (l > a.num_atcell) a.err[0] = 1;
    }
  }
  for (int na = 0; na < a.natom; na++) {
    if (a.natpri[na] == a.key_out && a.flag[na]) a.natpri[na] = a.key_inps;
    a.naps[na] = 0;
    if (a.natpri[na] == a.key_in || a.natpri[na] == a.key_inps) {
      a.naps[na] = ++iaps;
      if (iaps > a.num_ppcell) a.err[1] = 1;
    }
  }
  for (int na = 0; na < a.natom; na++) {
    a.napsd[na] = 0;
    if (a.natprid[na] == a.key_inps) {
      a.napsd[na] = ++iapsd;
      if (iapsd > a.num_ppcell_d) a.err[2] = 1;
    }
  }
}

// ordered compaction inside a work-group: position of this thread's entry among the flagged
// threads of the group, and the group's total (wave ballots, wave totals through LDS)
__device__ inline int block_rank(bool f, int *wtot, int &total) {
  const uint64_t m = __ballot(f);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int before = __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
  if (lane == 0) wtot[wv] = __popcll(m);
  __syncthreads();
  int off = 0;
  total = 0;
  for (int w = 0; w < TB / 64; w++) {
    if (w < wv) off += wtot[w];
    total += wtot[w];
  }
  __syncthreads();
  return off + before;
}

// projector list vectors of the in/inps atoms (pseudocalc8f.F90:954-1033)
__global__ void __launch_bounds__(TB) k_lists(ListArgs a) {
  __shared__ int wtot[TB / 64];
  const int na = blockIdx.x;
  if (a.natpri[na] != a.key_in && a.natpri[na] != a.key_inps) {
    if (threadIdx.x == 0) a.natinf[na] = 0;
    return;
  }
  const long col = (long)a.num_list * (a.naps[na] - 1);
  for (int l = threadIdx.x; l < a.num_list; l += TB)
    a.lstvec2[col + l] = a.lstx[col + l] = a.lsty[col + l] = a.lstz[col + l] = 0;
  __syncthreads();
  const CoarseBox b = coarse_box(a, na);
  const long np = 8L * a.npxmax * a.npymax * a.npzmax;
  int l0 = 0;
  for (long base = 0; base < np; base += TB) {
    const long p = base + threadIdx.x;
    int ix = 0, iy = 0, iz = 0, i = 0, j = 0, k = 0;
    const bool f = p < np && coarse_point(a, b, p, ix, iy, iz, i, j, k);
    int total;
    const int pos = l0 + block_rank(f, wtot, total);
    if (f && pos < a.num_list) {
      a.lstvec2[col + pos] = (k - 1) * a.ncpy * a.ncpx + (j - 1) * a.ncpx + i;
      a.lstx[col + pos] = ix;
      a.lsty[col + pos] = iy;
      a.lstz[col + pos] = iz;
    }
    l0 += total;
  }
  if (threadIdx.x == 0) {
    a.natinf[na] = l0;
    if (l0 > a.num_list) a.err[3] = 1;
  }
}

// dense list vectors (pseudocalc8f.F90:1036-1150): points inside radial(nradct), then the
// shell up to radial(nradct) * psctoff
__global__ void __launch_bounds__(TB) k_dense_lists(ListArgs a) {
  __shared__ int wtot[TB / 64];
  const int na = blockIdx.x;
  if (a.natprid[na] != a.key_inps) {
    if (threadIdx.x == 0) a.natinfd[na] = a.natinfd_vloc[na] = 0;
    return;
  }
  const long col = (long)a.num_list_d * (a.napsd[na] - 1);
  for (int l = threadIdx.x; l < a.num_list_d; l += TB)
    a.lstvecd2[col + l] = a.lstdx[col + l] = a.lstdy[col + l] = a.lstdz[col + l] = 0;
  __syncthreads();
  const DenseBox b = dense_box(a, na);
  const int m = a.nmesh;
  const long np = 8L * a.npxmax * a.npymax * a.npzmax * m * m * m;
  const int ndx = a.ncpx * m, ndy = a.ncpy * m;
  int l0 = 0;
  for (int pass = 1; pass <= 2; pass++) {
    for (long base = 0; base < np; base += TB) {
      const long p = base + threadIdx.x;
      int ix = 0, iy = 0, iz = 0, i = 0, j = 0, k = 0;
      const bool f = p < np && dense_point(a, b, p, ix, iy, iz, i, j, k) == pass;
      int total;
      const int pos = l0 + block_rank(f, wtot, total);
      if (f && pos < a.num_list_d) {
        a.lstvecd2[col + pos] = (k - 1) * ndy * ndx + (j - 1) * ndx + i;
        a.lstdx[col + pos] = ix;
        a.lstdy[col + pos] = iy;
        a.lstdz[col + pos] = iz;
      }
      l0 += total;
    }
    if (pass == 1 && threadIdx.x == 0) a.natinfd[na] = l0;
  }
  if (threadIdx.x == 0) {
    a.natinfd_vloc[na] = l0;
    if (l0 > a.num_list_d) a.err[4] = 1;
  }
}

// ============================================================================ pseudocalc_07
// One active atom's constants for the fine-grid projectors.
struct PPAtom {
  int offx, offy, offz;  // npat - npmesh * nat: fine index of box point 0 in vnlspd
  int sp, np, nq, mask, iaps, natinf, lcol;
  // the part of the fine work array (0-based array indices, start and extent per axis) where
  // the projectors can be nonzero: the atom's box cut to its sphere.  Every pass reads only
  // this part and treats the rest as the zeros it holds in the reference: skipping a zero term
  // leaves a sum unchanged bit for bit.
  int f0x, f0y, f0z, fwx, fwy, fwz;
  double ax, ay, az;  // atom minus the fine point it rounds to, plus half a fine step
  double rc, thr, dqq, tmp;
};

struct FineArgs {
  const PPAtom *at;
  const double *coef;
  int nqmx, nfiltyp;
  int nxspd, nyspd, nzspd, npxd, npyd, npzd;  // npxmax * npmesh ...
  double dpx, dpy, dpz, dpxyz, eps, filpp;
  double c0, c1, c1516, c1504, c0516;  // harmonic normalisations
  double *out;  // per atom slot: nprjmx planes of (2 nxspd)(2 nyspd)(2 nzspd)
  long plane;
  int nprjmx;
};

// the radial channels of each nprj (the cases of pseudocalc8f.F90:2589-3620): bit c set for
// coef channel c+1 (two l = 0, two l = 1, two l = 2); 0 for nprj values without a case
inline int pp_mask(int nprj) {
  switch (nprj) {
    case 1: return 0x01;
    case 2: return 0x03;
    case 4: return 0x05;
    case 5: return 0x07;
    case 7: return 0x0d;
    case 8: return 0x0f;
    case 9: return 0x15;
    case 10: return 0x17;
    case 12: return 0x1d;
    case 13: return 0x1f;
    case 14: return 0x35;
    case 15: return 0x37;
    case 17: return 0x3d;
    case 18: return 0x3f;
    default: return 0;
  }
}

// vnlspd of one atom slot (grid.y): every point of the fine work array, zero outside the
// atom's box and sphere (pseudocalc8f.F90:2583-3656)
__global__ void __launch_bounds__(256) k_fine_projectors(FineArgs a, int slot0) {
  const PPAtom &t = a.at[slot0 + blockIdx.y];
  double *out = a.out + (long)blockIdx.y * a.nprjmx * a.plane;
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= (long)t.fwx * t.fwy * t.fwz || t.np == 0) return;
  const int i0 = t.f0x + (int)(u % t.fwx), j0 = t.f0y + (int)((u / t.fwx) % t.fwy), k0 = t.f0z + (int)(u / ((long)t.fwx * t.fwy));
  const int sx = 2 * a.nxspd, sy = 2 * a.nyspd;
  const long q = i0 + (long)sx * (j0 + (long)sy * k0);
  const int kx = i0 - a.nxspd + 1, ky = j0 - a.nyspd + 1, kz = k0 - a.nzspd + 1;
  const int ix = kx - t.offx, iy = ky - t.offy, iz = kz - t.offz;
  const bool inbox = ix > -a.npxd && ix <= a.npxd && iy > -a.npyd && iy <= a.npyd && iz > -a.npzd && iz <= a.npzd;
  const double x = ix * a.dpx - t.ax, y = iy * a.dpy - t.ay, z = iz * a.dpz - t.az;
  const double r = sqrt(x * x + y * y + z * z);
  if (!(inbox && t.mask && r < t.thr)) {
    for (int l = 0; l < t.np; l++) out[(long)l * a.plane + q] = 0.0;
    return;
  }
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  double yl[10] = {0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0};
  double weight = 1.0;
  const long cs = a.nqmx + 1;
  const double *cf = a.coef + cs * 8 * (t.sp - 1);
  if (r > a.eps) {
    if (a.nfiltyp == 1) {
      const double ac = (r - t.rc) / a.filpp;
      weight = fabs(ac) < 1.0e2 ? 1.0 / (exp(ac) + 1.0) : (ac > 0.0 ? 0.0 : 1.0);
    }
    const double rin = 1.0 / r;
    yl[2] = rin * x;
    yl[3] = rin * y;
    yl[4] = rin * z;
    yl[5] = rin * rin * (x * x - y * y);
    yl[6] = -(rin * rin * z * x);
    yl[7] = rin * rin * (3.0 * z * z - r * r);
    yl[8] = -(rin * rin * y * z);
    yl[9] = rin * rin * x * y;
    const bool l1 = t.mask & 0x0c, l2 = t.mask & 0x30;
    for (int iq = 1; iq <= t.nq; iq++) {
      const double qqq = t.dqq * iq, qqq2 = qqq * qqq, qqqr = qqq * r;
      const double qqqrin = rin / qqq;
      double sn, cn;
      sincos(qqqr, &sn, &cn);
      const double j0 = sn * qqqrin;
      if (t.mask & 0x01) s[1] += cf[iq + cs * 1] * j0 * qqq2;
      if (t.mask & 0x02) s[2] += cf[iq + cs * 2] * j0 * qqq2;
      if (l1 || l2) {
        const double qqqr2in = qqqrin * qqqrin;
        const double j1 = (sn - qqqr * cn) * qqqr2in;
        if (t.mask & 0x04) s[3] += cf[iq + cs * 3] * j1 * qqq2;
        if (t.mask & 0x08) s[4] += cf[iq + cs * 4] * j1 * qqq2;
        if (l2) {
          const double j2 = ((3.0 - qqqr * qqqr) * sn - 3.0 * qqqr * cn) * (qqqr2in * qqqrin);
          if (t.mask & 0x10) s[5] += cf[iq + cs * 5] * j2 * qqq2;
          if (t.mask & 0x20) s[6] += cf[iq + cs * 6] * j2 * qqq2;
        }
      }
    }
  } else {
    for (int iq = 1; iq <= t.nq; iq++) {
      const double qqq = t.dqq * iq;
      if (t.mask & 0x01) s[1] += cf[iq + cs * 1] * qqq * qqq;
      if (t.mask & 0x02) s[2] += cf[iq + cs * 2] * qqq * qqq;
    }
  }
  // projector order: the channels in order, each with its 2l+1 real harmonics
  int la = 0;
#pragma unroll
  for (int c = 1; c <= 6; c++) {
    if (!(t.mask & (1 << (c - 1)))) continue;
    const double sc = s[c] * t.tmp * a.dpxyz;
    if (c <= 2) {
      out[(long)la++ * a.plane + q] = sc * a.c0 * yl[1] * weight;
    } else if (c <= 4) {
      out[(long)la++ * a.plane + q] = sc * a.c1 * yl[2] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c1 * yl[3] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c1 * yl[4] * weight;
    } else {
      out[(long)la++ * a.plane + q] = sc * a.c1516 * yl[5] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c1504 * yl[6] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c0516 * yl[7] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c1504 * yl[8] * weight;
      out[(long)la++ * a.plane + q] = sc * a.c1504 * yl[9] * weight;
    }
  }
  for (; la < t.np; la++) out[(long)la * a.plane + q] = 0.0;
}

struct FilterArgs {
  const PPAtom *at;
  const double *w;  // wxyz from its lower bound -nfdg*npmesh
  int js, je, wlo, npmesh, half;
  int nxspd, nyspd, nzspd, npxmax, npymax, npzmax, nprjmx;
  long plane, plane1, plane2;  // per projector: fine array, after z, after y
  const double *fine;
  double *t1, *t2;
};

// z pass: t1(ix, iy, iz) over fine x, y and coarse z (pseudocalc8f.F90:3662-3671), on the
// columns of the atom's fine part
__global__ void k_filter_z(FilterArgs a, int slot0) {
  const PPAtom &t = a.at[slot0 + blockIdx.y];
  const int la = blockIdx.z;
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cz = 2 * a.npzmax;
  if (la >= t.np || u >= (long)t.fwx * t.fwy * cz) return;
  const int sx = 2 * a.nxspd, sy = 2 * a.nyspd;
  const int i0 = t.f0x + (int)(u % t.fwx), j0 = t.f0y + (int)((u / t.fwx) % t.fwy);
  const int izc = (int)(u / ((long)t.fwx * t.fwy)), iz = izc - a.npzmax + 1;
  const long ixy = i0 + (long)sx * j0;
  const double *src = a.fine + ((long)blockIdx.y * a.nprjmx + la) * a.plane + ixy;
  const int kkz = iz * a.npmesh - a.half;
  double acc = 0.0;
  for (int jz = a.js; jz <= a.je; jz++) {
    const int k0 = kkz + jz + a.nzspd - 1;
    if (k0 >= t.f0z && k0 < t.f0z + t.fwz) acc = acc + src[(long)k0 * sx * sy] * a.w[jz - a.wlo];
  }
  a.t1[((long)blockIdx.y * a.nprjmx + la) * a.plane1 + ixy + (long)sx * sy * izc] = acc;
}

// y pass: t2(ix, iy, iz) over fine x and coarse y, z (:3673-3682), on the atom's fine x range
__global__ void k_filter_y(FilterArgs a, int slot0) {
  const PPAtom &t = a.at[slot0 + blockIdx.y];
  const int la = blockIdx.z;
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cy = 2 * a.npymax, cz = 2 * a.npzmax;
  if (la >= t.np || u >= (long)t.fwx * cy * cz) return;
  const int sx = 2 * a.nxspd, sy = 2 * a.nyspd;
  const int i0 = t.f0x + (int)(u % t.fwx), iyc = (int)((u / t.fwx) % cy), izc = (int)(u / ((long)t.fwx * cy));
  const int iy = iyc - a.npymax + 1;
  const double *src = a.t1 + ((long)blockIdx.y * a.nprjmx + la) * a.plane1 + (long)izc * sx * sy + i0;
  const int kky = iy * a.npmesh - a.half;
  double acc = 0.0;
  for (int jy = a.js; jy <= a.je; jy++) {
    const int j0 = kky + jy + a.nyspd - 1;
    if (j0 >= t.f0y && j0 < t.f0y + t.fwy) acc = acc + src[(long)j0 * sx] * a.w[jy - a.wlo];
  }
  a.t2[((long)blockIdx.y * a.nprjmx + la) * a.plane2 + i0 + (long)sx * (iyc + (long)cy * izc)] = acc;
}

// x pass at the atom's list points, into vnlocp(:, la, iaps) (:3684-3698)
__global__ void k_filter_x_gather(FilterArgs a, int slot0, const int *lstx, const int *lsty, const int *lstz,
                                  double *vnlocp, int num_list) {
  const PPAtom &t = a.at[slot0 + blockIdx.y];
  const int la = blockIdx.z;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (la >= t.np || p >= t.natinf) return;
  const int sx = 2 * a.nxspd, cy = 2 * a.npymax;
  const long li = (long)t.lcol + p;
  const int ix = lstx[li], iy = lsty[li], iz = lstz[li];
  const double *src = a.t2 + ((long)blockIdx.y * a.nprjmx + la) * a.plane2 +
                      ((long)(iz + a.npzmax - 1) * cy + (iy + a.npymax - 1)) * sx;
  const int kkx = ix * a.npmesh - a.half;
  double acc = 0.0;
  for (int jx = a.js; jx <= a.je; jx++) {
    const int i0 = kkx + jx + a.nxspd - 1;
    if (i0 >= t.f0x && i0 < t.f0x + t.fwx) acc = acc + src[i0] * a.w[jx - a.wlo];
  }
  vnlocp[(long)num_list * (la + (long)a.nprjmx * (t.iaps - 1)) + p] = acc;
}

}  // namespace

extern "C" {

int kukan_pseudocalc_01a(int nfdg, int npmesh, double *wxyz) {
  API_BEGIN
  if (nfdg < 1 || nfdg > 8) throw Error(1, "error in pseudo. nfdg should be between 1 and 8.");
  if (npmesh < 1) throw Error(1, "pseudocalc_01a: npmesh must be positive");
  const int nw = 2 * nfdg * npmesh;
  Arr w(wxyz, (size_t)nw * D, Arr::Out);
  hipLaunchKernelGGL(k_dg_weights, dim3((nw + 255) / 256), dim3(256), 0, lib().stream, nfdg, npmesh, w.d());
  HIPCK(hipGetLastError());
  w.finish();
  API_END
}

int kukan_pseudocalc_01b(int nperi, int natom, int num_spe, int nradmx, int nmesh, int nfdg, int ndisp,
                         int num_atcell, int num_ppcell, int num_ppcell_d, int num_list, int num_list_d,
                         int nxmax, int nymax, int nzmax, int ncpx, int ncpy, int ncpz, int npxmax, int npymax,
                         int npzmax, int key_natpri_in, int key_natpri_inps, int key_natpri_out, double psctoff,
                         double xmax, double ymax, double zmax, double dx, double dy, double dz,
                         const int *indspe, const int *nradct, const double *radial, const double *atx,
                         const double *aty, const double *atz, int *natx, int *naty, int *natz, int *ndatx,
                         int *ndaty, int *ndatz, int *lstx, int *lsty, int *lstz, int *lstdx, int *lstdy,
                         int *lstdz, int *natinf, int *natinfd, int *natinfd_vloc, int *lstvec2, int *lstvecd2,
                         int *natpri, int *natprid, int *natpri_inf, int *naps, int *napsd) {
  (void)ndisp;
  API_BEGIN
  if (natom < 1 || num_spe < 1 || nmesh < 1 || nfdg < 0 || npxmax < 1 || npymax < 1 || npzmax < 1 ||
      ncpx < 1 || ncpy < 1 || ncpz < 1 || nradmx < 1)
    throw Error(1, "pseudocalc_01b: bad sizes");
  if (ncpx != 2 * nxmax || ncpy != 2 * nymax || ncpz != 2 * nzmax)
    throw Error(1, "pseudocalc_01b: one rank holds the whole grid (ncpx = 2 nxmax ...)");
  {
    const std::vector<int> isp = host_ints(indspe, natom), nrc = host_ints(nradct, num_spe);
    for (int na = 0; na < natom; na++)
      if (isp[na] < 1 || isp[na] > num_spe || nrc[isp[na] - 1] < 1 || nrc[isp[na] - 1] > nradmx)
        throw Error(1, "pseudocalc_01b: indspe / nradct out of range");
  }
  hipStream_t s = lib().stream;
  const size_t an = (size_t)natom * I;
  Arr Ai(indspe, an, Arr::In), Anr(nradct, (size_t)num_spe * I, Arr::In);
  Arr Arad(radial, (size_t)nradmx * num_spe * D, Arr::In);
  Arr Ax(atx, (size_t)natom * D, Arr::In), Ay(aty, (size_t)natom * D, Arr::In), Az(atz, (size_t)natom * D, Arr::In);
  Arr Onx(natx, an, Arr::Out), Ony(naty, an, Arr::Out), Onz(natz, an, Arr::Out);
  Arr Odx(ndatx, an, Arr::Out), Ody(ndaty, an, Arr::Out), Odz(ndatz, an, Arr::Out);
  Arr Opri(natpri, an, Arr::Out), Oprid(natprid, an, Arr::Out), Oinf(natpri_inf, an, Arr::Out);
  Arr Onaps(naps, an, Arr::Out), Onapsd(napsd, an, Arr::Out);
  Arr Onatinf(natinf, an, Arr::Out), Onatinfd(natinfd, an, Arr::Out), Onatinfdv(natinfd_vloc, an, Arr::Out);
  // list columns of atoms that are not listed keep their contents
  const size_t lb = (size_t)num_list * num_ppcell * I, lbd = (size_t)num_list_d * num_ppcell_d * I;
  Arr Ol(lstvec2, lb, Arr::InOut), Olx(lstx, lb, Arr::InOut), Oly(lsty, lb, Arr::InOut), Olz(lstz, lb, Arr::InOut);
  Arr Od(lstvecd2, lbd, Arr::InOut), Odlx(lstdx, lbd, Arr::InOut), Odly(lstdy, lbd, Arr::InOut),
      Odlz(lstdz, lbd, Arr::InOut);
  int *flag = (int *)scratch("pp.flag", an + 8 * I);
  int *err = flag + natom;
  HIPCK(hipMemsetAsync(err, 0, 8 * I, s));
  ListArgs a{natom, nmesh, nfdg, nperi, nradmx, nxmax, nymax, nzmax, ncpx, ncpy, ncpz, npxmax, npymax, npzmax,
             num_list, num_list_d, num_atcell, num_ppcell, num_ppcell_d, key_natpri_in, key_natpri_inps,
             key_natpri_out, psctoff, xmax, ymax, zmax, dx, dy, dz, Ai.i(), Anr.i(), Arad.d(), Ax.d(), Ay.d(),
             Az.d(), Onx.i(), Ony.i(), Onz.i(), Odx.i(), Ody.i(), Odz.i(), Opri.i(), Oprid.i(), Oinf.i(),
             Onaps.i(), Onapsd.i(), Onatinf.i(), Onatinfd.i(), Onatinfdv.i(), Ol.i(), Olx.i(), Oly.i(), Olz.i(),
             Od.i(), Odlx.i(), Odly.i(), Odlz.i(), flag, err};
  {
    Timed t("pp_lists", 0.0);
    hipLaunchKernelGGL(k_atom_info, dim3((natom + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_sphere_flags, dim3(natom), dim3(TB), 0, s, a);
    hipLaunchKernelGGL(k_counters, dim3(1), dim3(1), 0, s, a);
    HIPCK(hipGetLastError());
  }
  int herr[5] = {0, 0, 0, 0, 0};
  HIPCK(hipMemcpyAsync(herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  sync();
  if (herr[0]) throw Error(1, "pseudocalc_01b: num_atcell is too small.");
  if (herr[1]) throw Error(1, "error! num_ppcell is too small.");
  if (herr[2]) throw Error(1, "error! num_ppcell_d is too small.");
  {
    Timed t("pp_lists", 0.0);
    hipLaunchKernelGGL(k_lists, dim3(natom), dim3(TB), 0, s, a);
    hipLaunchKernelGGL(k_dense_lists, dim3(natom), dim3(TB), 0, s, a);
    HIPCK(hipGetLastError());
  }
  HIPCK(hipMemcpyAsync(herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  sync();
  if (herr[3]) throw Error(1, "error! num_list is too small.");
  if (herr[4]) throw Error(1, "error! num_list_d is too small.");
  for (Arr *o : {&Onx, &Ony, &Onz, &Odx, &Ody, &Odz, &Opri, &Oprid, &Oinf, &Onaps, &Onapsd, &Onatinf, &Onatinfd,
                 &Onatinfdv, &Ol, &Olx, &Oly, &Olz, &Od, &Odlx, &Odly, &Odlz})
    o->finish();
  API_END
}

int kukan_pseudocalc_07(int natom, int num_spe, int npmesh, int nprjmx, int nfdg, int nfiltyp, int num_list,
                        int num_ppcell, int nradmx, int nqmx, int nxmax, int nymax, int nzmax, int npxmax,
                        int npymax, int npzmax, int nxspd, int nyspd, int nzspd, int key_natpri_in,
                        int key_natpri_inps, double eps, double psftrad, double psctoff, double filpp, double dx,
                        double dy, double dz, const int *indspe, const int *nprj, const int *nradct,
                        const int *natpri, const int *naps, const int *natinf, const int *nqct, const int *natx,
                        const int *naty, const int *natz, const int *lstx, const int *lsty, const int *lstz,
                        const double *wxyz, const double *radial, const double *coef, const double *atx,
                        const double *aty, const double *atz, double *vnlocp, double *vtmp_pp, double *vtmp_pp1,
                        double *vtmp_pp2, double *vnlspd) {
  // the work arrays of the reference are not used: the passes run in device workspace
  (void)vtmp_pp;
  (void)vtmp_pp1;
  (void)vtmp_pp2;
  (void)vnlspd;
  API_BEGIN
  if (nfdg < 1 || nfdg > 8 || npmesh < 1) throw Error(1, "pseudocalc_07: nfdg must be 1..8, npmesh >= 1");
  if (nprjmx < 1 || nprjmx > 18) throw Error(1, "pseudocalc_07: nprjmx must be 1..18");
  const int half = (npmesh - 1) / 2;
  if (nxspd < npmesh * (npxmax + nfdg) - (npmesh + 1) / 2 || nyspd < npmesh * (npymax + nfdg) - (npmesh + 1) / 2 ||
      nzspd < npmesh * (npzmax + nfdg) - (npmesh + 1) / 2)
    throw Error(1, "pseudocalc_07: nxspd/nyspd/nzspd smaller than the double-grid filter reads");
  hipStream_t s = lib().stream;
  const double pi = std::acos(-1.0);
  const double dpx = dx / npmesh, dpy = dy / npmesh, dpz = dz / npmesh;
  const double twopicbin = 1.0 / ((2.0 * pi) * (2.0 * pi) * (2.0 * pi)), fourpi = 4.0 * pi;
  const std::vector<int> isp = host_ints(indspe, natom), pri = host_ints(natpri, natom), ap = host_ints(naps, natom),
                         ninf = host_ints(natinf, natom), nx = host_ints(natx, natom), ny = host_ints(naty, natom),
                         nz = host_ints(natz, natom), npj = host_ints(nprj, num_spe), nqc = host_ints(nqct, num_spe),
                         nrc = host_ints(nradct, num_spe);
  const std::vector<double> rad = host_doubles(radial, (size_t)nradmx * num_spe), hx = host_doubles(atx, natom),
                            hy = host_doubles(aty, natom), hz = host_doubles(atz, natom);
  std::vector<PPAtom> at;
  for (int na = 0; na < natom; na++) {
    if (pri[na] != key_natpri_in && pri[na] != key_natpri_inps) continue;
    const int sp = isp[na];
    if (sp < 1 || sp > num_spe) throw Error(1, "pseudocalc_07: indspe out of range");
    if (ap[na] < 1 || ap[na] > num_ppcell) throw Error(1, "pseudocalc_07: naps out of range");
    if (ninf[na] < 0 || ninf[na] > num_list) throw Error(1, "pseudocalc_07: natinf out of range");
    if (nrc[sp - 1] < 1 || nrc[sp - 1] > nradmx) throw Error(1, "pseudocalc_07: nradct out of range");
    const int np = npj[sp - 1];
    if (np > nprjmx) throw Error(1, "pseudocalc_07: nprj exceeds nprjmx");
    if (nqc[sp - 1] > nqmx) throw Error(1, "pseudocalc_07: nqct exceeds nqmx");
    PPAtom t{};
    const int npx = (int)((hx[na] + nxmax * dx + 0.5 * dpx) / dpx) - nxmax * npmesh;
    const int npy = (int)((hy[na] + nymax * dy + 0.5 * dpy) / dpy) - nymax * npmesh;
    const int npz = (int)((hz[na] + nzmax * dz + 0.5 * dpz) / dpz) - nzmax * npmesh;
    t.offx = npx - npmesh * nx[na];
    t.offy = npy - npmesh * ny[na];
    t.offz = npz - npmesh * nz[na];
    // the fine box must lie inside the work array (pseudocalc8f.F90:227-229 sizes it so)
    if (std::abs(t.offx) > nxspd - npxmax * npmesh || std::abs(t.offy) > nyspd - npymax * npmesh ||
        std::abs(t.offz) > nzspd - npzmax * npmesh)
      throw Error(1, "pseudocalc_07: atom too far from its coarse point (natx.. not from pseudocalc_01b?)");
    t.sp = sp;
    t.mask = pp_mask(np);
    t.np = np;  // projectors without a case stay zero, as in the reference
    t.nq = nqc[sp - 1];
    t.iaps = ap[na];
    t.natinf = ninf[na];
    t.lcol = num_list * (ap[na] - 1);
    if (nqc[sp - 1] < 0) throw Error(1, "pseudocalc_07: nqct negative");
    t.ax = hx[na] - (npx * dpx - 0.5 * dpx);
    t.ay = hy[na] - (npy * dpy - 0.5 * dpy);
    t.az = hz[na] - (npz * dpz - 0.5 * dpz);
    t.rc = rad[(size_t)(nrc[sp - 1] - 1) + (size_t)nradmx * (sp - 1)];
    t.thr = t.rc * psctoff;
    t.dqq = pi / (t.rc * psftrad);
    t.tmp = fourpi * t.dqq * twopicbin;
    // the fine part: box points within the sphere's bounding box (one point of margin)
    auto part = [&](double c, double h, int nbox, int off, int nspd, int &f0, int &fw) {
      const int lo = std::max(-nbox + 1, (int)std::floor((c - t.thr) / h) - 1);
      const int hi = std::min(nbox, (int)std::ceil((c + t.thr) / h) + 1);
      f0 = lo + off + nspd - 1;
      fw = std::max(0, hi - lo + 1);
    };
    part(t.ax, dpx, npxmax * npmesh, t.offx, nxspd, t.f0x, t.fwx);
    part(t.ay, dpy, npymax * npmesh, t.offy, nyspd, t.f0y, t.fwy);
    part(t.az, dpz, npzmax * npmesh, t.offz, nzspd, t.f0z, t.fwz);
    at.push_back(t);
  }
  const size_t vb = (size_t)num_list * nprjmx * num_ppcell * D;
  Arr V(vnlocp, vb, Arr::Out);
  HIPCK(hipMemsetAsync(V.d(), 0, vb, s));
  const int nact = (int)at.size();
  if (nact > 0) {
    const int nw = 2 * nfdg * npmesh;
    Arr W(wxyz, (size_t)nw * D, Arr::In);
    Arr C(coef, (size_t)(nqmx + 1) * 8 * num_spe * D, Arr::In);
    {
      // the read-out indexes the filtered box through lstx/lsty/lstz
      const size_t nl = (size_t)num_list * num_ppcell;
      const std::vector<int> hlx = host_ints(lstx, nl), hly = host_ints(lsty, nl), hlz = host_ints(lstz, nl);
      for (const PPAtom &t : at)
        for (int p = 0; p < t.natinf; p++) {
          const size_t li = (size_t)t.lcol + p;
          if (hlx[li] <= -npxmax || hlx[li] > npxmax || hly[li] <= -npymax || hly[li] > npymax ||
              hlz[li] <= -npzmax || hlz[li] > npzmax)
            throw Error(1, "pseudocalc_07: lstx/lsty/lstz outside the atom's box");
        }
    }
    Arr Lx(lstx, (size_t)num_list * num_ppcell * I, Arr::In), Ly(lsty, (size_t)num_list * num_ppcell * I, Arr::In),
        Lz(lstz, (size_t)num_list * num_ppcell * I, Arr::In);
    PPAtom *atd = (PPAtom *)scratch("pp.atoms", sizeof(PPAtom) * at.size());
    HIPCK(hipMemcpyAsync(atd, at.data(), sizeof(PPAtom) * at.size(), hipMemcpyHostToDevice, s));
    const long plane = 8L * nxspd * nyspd * nzspd;
    const long plane1 = 8L * nxspd * nyspd * npzmax, plane2 = 8L * nxspd * npymax * npzmax;
    // atoms per batch: the fine arrays and both filter intermediates of a batch in ~1 GiB
    const long per_atom = (long)nprjmx * (plane + plane1 + plane2) * (long)D;
    const int batch = (int)std::max(1L, std::min<long>(nact, (1L << 30) / std::max(1L, per_atom)));
    double *fine = (double *)scratch("pp.fine", (size_t)batch * nprjmx * plane * D);
    double *t1 = (double *)scratch("pp.t1", (size_t)batch * nprjmx * plane1 * D);
    double *t2 = (double *)scratch("pp.t2", (size_t)batch * nprjmx * plane2 * D);
    FineArgs fa{atd, C.d(), nqmx, nfiltyp, nxspd, nyspd, nzspd, npxmax * npmesh, npymax * npmesh,
                npzmax * npmesh, dpx, dpy, dpz, dpx * dpy * dpz, eps, filpp,
                std::sqrt(1.0 / 4.0 / pi), std::sqrt(3.0 / 4.0 / pi), std::sqrt(15.0 / 16.0 / pi),
                std::sqrt(15.0 / 4.0 / pi), std::sqrt(5.0 / 16.0 / pi), fine, plane, nprjmx};
    const int wlo = -nfdg * npmesh;
    FilterArgs ga{atd, W.d(), (npmesh % 2 == 1) ? wlo + 1 : wlo, nfdg * npmesh - 1, wlo, npmesh, half, nxspd,
                  nyspd, nzspd, npxmax, npymax, npzmax, nprjmx, plane, plane1, plane2, fine, t1, t2};
    int maxinf = 0;
    long maxf = 0, maxz = 0, maxy = 0;
    for (const PPAtom &t : at) {
      maxinf = std::max(maxinf, t.natinf);
      maxf = std::max(maxf, (long)t.fwx * t.fwy * t.fwz);
      maxz = std::max(maxz, (long)t.fwx * t.fwy * 2 * npzmax);
      maxy = std::max(maxy, (long)t.fwx * 2 * npymax * 2 * npzmax);
    }
    for (int a0 = 0; a0 < nact; a0 += batch) {
      const int nb = std::min(batch, nact - a0);  // <= 65535 (grid.y): batch holds ~1 GiB
      {
        Timed t("pp_fine", 0.0);
        if (maxf > 0)
          hipLaunchKernelGGL(k_fine_projectors, dim3((unsigned)((maxf + 255) / 256), nb), dim3(256), 0, s, fa, a0);
      }
      {
        Timed t("pp_filter", 0.0);
        if (maxz > 0)
          hipLaunchKernelGGL(k_filter_z, dim3((unsigned)((maxz + 255) / 256), nb, nprjmx), dim3(256), 0, s, ga, a0);
        if (maxy > 0)
          hipLaunchKernelGGL(k_filter_y, dim3((unsigned)((maxy + 255) / 256), nb, nprjmx), dim3(256), 0, s, ga, a0);
        if (maxinf > 0)
          hipLaunchKernelGGL(k_filter_x_gather, dim3((unsigned)((maxinf + 255) / 256), nb, nprjmx), dim3(256), 0, s,
                             ga, a0, Lx.i(), Ly.i(), Lz.i(), V.d(), num_list);
      }
      HIPCK(hipGetLastError());
    }
    sync();  // the atom table's host copy goes out of scope
  }
  V.finish();
  API_END
}

}  // extern "C"
