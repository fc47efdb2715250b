// kernels.hip — device kernels of libkukan_hip.so for gfx950 (CDNA4, wave64).
//
// Layouts follow the reference's Fortran arrays: a band is ncpx*ncpy*ncpz doubles, x
// fastest; bands are contiguous columns (svecre(:,:,:,l)).  Halos are never materialised
// for the solver paths: stencil kernels read neighbours with a periodic wrap (or zero, for
// isolated directions) straight from the unhaloed band, which replaces the reference's
// pack/copy halo exchange (overlap_finitedifference8f.F90:97-137) on one GPU.
#include <hipcub/hipcub.hpp>

#include "kukan_internal.hpp"

namespace kukan {
namespace {

constexpr int WAVE = 64;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
  return v;
}

// Block-wide sum for blockDim.x*blockDim.y*blockDim.z == NT threads; result valid in thread 0.
template <int NT>
__device__ inline double block_sum(double v, double *red) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_sum(v);
  if ((t & (WAVE - 1)) == 0) red[t / WAVE] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (int w = 0; w < NT / WAVE; w++) s += red[w];
  __syncthreads();
  return s;
}

// Block-wide max of non-negative values (same shape rules as block_sum); result in thread 0.
template <int NT>
__device__ inline double block_max(double v, double *red) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, WAVE));
  if ((t & (WAVE - 1)) == 0) red[t / WAVE] = v;
  __syncthreads();
  double m = 0.0;
  if (t == 0)
    for (int w = 0; w < NT / WAVE; w++) m = fmax(m, red[w]);
  __syncthreads();
  return m;
}

// ----------------------------------------------------------------------------- stencil
struct SP {
  int nx, ny, nz, px, py, pz, halo, nfin;
  long ex, ey;
  const double *in;
  long ld_in;
  double *out;
  long ld_out;
  const double *veff;
  double cx[9], cy[9], cz[9], c0;
  int accumulate;
  const int *skip;
  const int *row_ptr;
  const int2 *ent;
  const double *vnl;
  const double *coef;
  long ld_coef;
  int num_list, nprjmx;
  const int *iprj_slot;
  int zc, nzch;
  double *dotp;
  long dot_ld;
};

__device__ inline double val(const SP &p, const double *base, int x, int y, int z) {
  if (p.halo) return base[(long)(x + p.nfin) + p.ex * ((long)(y + p.nfin) + p.ey * (long)(z + p.nfin))];
  if (x < 0 || x >= p.nx) {
    if (!p.px) return 0.0;
    x = x < 0 ? x + p.nx : x - p.nx;
  }
  if (y < 0 || y >= p.ny) {
    if (!p.py) return 0.0;
    y = y < 0 ? y + p.ny : y - p.ny;
  }
  if (z < 0 || z >= p.nz) {
    if (!p.pz) return 0.0;
    z = z < 0 ? z + p.nz : z - p.nz;
  }
  return base[(long)x + (long)p.nx * ((long)y + (long)p.ny * z)];
}

// One thread per (x,y) column marching over a z-chunk; the z-neighbours live in a register
// queue, x/y neighbours come through the vector L1.  Block 64 x 4 (one wavefront per x-row).
template <int NF>
__global__ __launch_bounds__(256) void k_stencil(SP p) {
  __shared__ double red[4];
  if (p.skip && *p.skip) return;  // uniform: a device-side loop has finished
  const int nxb = (p.nx + 63) / 64;
  const int x = (blockIdx.x % nxb) * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int zch = blockIdx.x / nxb, b = blockIdx.z;
  const int z0 = zch * p.zc;
  const int z1 = min(z0 + p.zc, p.nz);
  const bool act = x < p.nx && y < p.ny;
  const double *in = p.in + (long)b * p.ld_in;
  double *out = p.out + (long)b * p.ld_out;
  double q[2 * NF + 1];
  double dsum = 0.0;
  if (act) {
#pragma unroll
    for (int m = 0; m < 2 * NF; m++) q[m + 1] = val(p, in, x, y, z0 - NF + m);
    for (int z = z0; z < z1; z++) {
#pragma unroll
      for (int m = 0; m < 2 * NF; m++) q[m] = q[m + 1];
      q[2 * NF] = val(p, in, x, y, z + NF);
      const long r = (long)x + (long)p.nx * ((long)y + (long)p.ny * z);
      const double c = q[NF];
      double s = (p.c0 + (p.veff ? p.veff[r] : 0.0)) * c;
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cx[m] * (val(p, in, x - m, y, z) + val(p, in, x + m, y, z));
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cy[m] * (val(p, in, x, y - m, z) + val(p, in, x, y + m, z));
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cz[m] * (q[NF - m] + q[NF + m]);
      if (p.row_ptr) {
        const double *cf = p.coef + (long)b * p.ld_coef;
        const int e1 = p.row_ptr[r + 1];
        for (int e = p.row_ptr[r]; e < e1; e++) {
          const int2 en = p.ent[e];
          const int ip = p.iprj_slot[en.x];
          const double *v = p.vnl + en.y + (long)p.num_list * p.nprjmx * en.x;
          const double *cc = cf + (long)p.nprjmx * en.x;
          for (int j = 0; j < ip; j++) s += cc[j] * v[(long)p.num_list * j];
        }
      }
      if (p.dotp) dsum += c * s;
      out[r] = p.accumulate ? out[r] + s : s;
    }
  }
  if (p.dotp) {
    const double t = block_sum<256>(dsum, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      const long part = blockIdx.x + (long)gridDim.x * blockIdx.y;
      const long nparts = (long)gridDim.x * gridDim.y;
      p.dotp[(long)b * (p.dot_ld ? p.dot_ld : nparts) + part] = t;
    }
  }
}

// ----------------------------------------------------------------------------- LDS stencil
// z-marching tile kernel for interior (unhaloed) bands, nx <= 256.
//   * a block owns full x-rows: a tile of nx x TY points, marching over a z-range; each
//     thread owns P points of the tile and keeps their z-columns (2*NF+1 planes) in
//     registers, so every band value is read from HBM once per z-range;
//   * the current plane of the tile plus NF halo rows above and below sit in LDS
//     (x wraps inside the row), giving the x and y neighbours;
//   * blocks are mapped XCD-aware: all tiles and z-ranges of a band run on one XCD
//     (block id mod 8), so the halo rows and veff/projector reads hit that XCD's L2;
//   * MODE 1 reads sign*(h - eps_b s) instead of a band (fused residual for K r); MODE 2 gives
//     H (ca_b x + cb_b y) by linearity, as ca_b H x + cb_b (H y) with H y read (hin), so the
//     combination is never stored and the stencil loads only x; it also sums <y|out>.
//     MODE 5 is MODE 2 without y: for a symmetric H, <y|H p> = ca_b <H y|x> + cb_b <y|H y>,
//     so the block sums <ca_b x|out> and <ca_b x|H y> from x and H y alone and y is not read.
//   * MODE 6 (radius 1) is the block CG's band update fused with the next residual
//     preconditioner: every value it loads is formed as xn = ca_b x + cb_b y (the new band),
//     un = ha_b h + hb_b w (its H) and r = sign (un - eps_b xn); out = K r, and the centre's
//     xn, un go to out_x, out_h (other buffers than x, h: neighbouring blocks still read those),
//     with partials <r|K r> (dotp), <xn|xn> (dotp2), <xn|un> (dotp3) and max |K r| (omax).
struct SL {
  int nx, ny, nz, px, py, pz, TY, ntile, nzch, zc, nbands;
  const double *in, *in2;
  long ld_in;
  const double *eps;
  double sign;
  double *out;
  long ld_out;
  const double *veff;
  double cx[9], cy[9], cz[9], c0;
  int accumulate;
  const int *skip;
  const int *row_ptr;
  const int2 *ent;
  const double *vnl;
  const double *coef;
  long ld_coef;
  int num_list, nprjmx;
  const int *iprj_slot;
  double *dotp;  // per band: ntile * nzch partials of <c|out> (c = centre input value)
  long dot_ld;   // band stride of dotp
  unsigned long long *omax;  // optional: per band max |out| as FP64 bits (atomic max)
  const double *ca, *cb;     // MODE 2: the input is ca_b in + cb_b in2
  double *dotp2;             // MODE 2: partials of <in2|out> (in2 at the centre), as dotp
  const double *hin;         // MODE 2: H in2 (same band stride as in)
  unsigned long long *cmax;  // optional: per band max |c| (centre input value) as FP64 bits
  double *dotp3;             // optional: partials of <c|c>, as dotp
  const double *hin2, *ha, *hb;  // MODE 6: w, and the coefficients of un
  double *out_x, *out_h;         // MODE 6: xn and un at the centre
};

// Block = TY full x-rows of one band, marching over a z-range.  The 2 NF + 1 values of each
// point's z-column live in registers; plane z (TY rows plus NF halo rows on either side, each
// row padded by NF wrapped or zero columns) is staged in LDS, so every x and y neighbour is an
// unconditional LDS read.  Loads of plane z + NF + 1, of plane z + 1's halo rows and of
// V_eff(z + 1) are issued before plane z is computed (one plane of prefetch), and the barriers
// wait only for LDS, so HBM latency overlaps the arithmetic.
// One work-group's share: band b, y-tile `tile`, z-range `zch` (all arguments uniform).
template <int NF, int P, int MODE>
__device__ inline void stencil_block(const SL &p, int b, int tile, int zch, double *lds,
                                     double *red) {
  constexpr bool RES = MODE == 1, COMB = MODE == 2 || MODE == 5, SYM = MODE == 5, UPD = MODE == 6;
  static_assert(!UPD || NF == 1, "the fused update is radius 1");
  // MODE 3: plain, also the partials of <c|c> (no max); MODE 4: plain, also the maxima of |c|
  // and |out| (no partials) - each keeps the plain kernel's accumulator count
  constexpr bool NORMS = MODE == 3, MAXES = MODE == 4;
  constexpr int H = NF;  // halo values per thread: 2 NF rows x nx <= 512 NF for nx <= 256
  const int nx = p.nx, ny = p.ny, nz = p.nz, TY = p.TY, W = nx + 2 * NF;
  const long plane = (long)nx * ny;
  const int per = p.ntile * p.nzch;
  const int y0 = tile * TY, z0 = zch * p.zc, z1 = min(nz, z0 + p.zc);
  const double *in = p.in + (long)b * p.ld_in;
  const double *in2 = (RES || UPD || (COMB && !SYM)) ? p.in2 + (long)b * p.ld_in : nullptr;
  const double *hin = (COMB || UPD) ? p.hin + (long)b * p.ld_in : nullptr;
  const double *hin2 = UPD ? p.hin2 + (long)b * p.ld_in : nullptr;
  const double e = (RES || UPD) ? p.eps[b] : 0.0;
  const double fa = (COMB || UPD) ? p.ca[b] : 0.0, fb = (COMB || UPD) ? p.cb[b] : 0.0;
  const double ga = UPD ? p.ha[b] : 0.0, gb = UPD ? p.hb[b] : 0.0;
  double *out = p.out + (long)b * p.ld_out;
  const int t = threadIdx.x;
  // y of a row index relative to the grid, wrapped once when periodic; -1 = zero row
  auto wrap_y = [&](int gy) -> int {
    if (gy >= 0 && gy < ny) return gy;
    if (!p.py) return -1;
    gy = gy < 0 ? gy + ny : gy - ny;
    return (gy >= 0 && gy < ny) ? gy : -1;
  };
  // z plane offset, -1 = zero plane (uniform across the block)
  auto zoff = [&](int z) -> long {
    if (z < 0 || z >= nz) {
      if (!p.pz) return -1;
      z = z < 0 ? z + nz : z - nz;
    }
    return (long)z * plane;
  };
  auto ld = [&](long off) -> double {
    if (RES) return p.sign * (in[off] - e * in2[off]);
    if (UPD) {
      const double xn = fa * in[off] + fb * in2[off], un = ga * hin[off] + gb * hin2[off];
      return p.sign * (un - e * xn);
    }
    return in[off];
  };
  // MODE 6: the same, also returning xn and un
  auto ld3 = [&](long off, double &xn, double &un) -> double {
    xn = fa * in[off] + fb * in2[off];
    un = ga * hin[off] + gb * hin2[off];
    return p.sign * (un - e * xn);
  };
  // tile points
  int oxy[P], lcol[P];
  bool lv[P], sv[P];
  int xq[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    const int idx = t + 512 * q;
    const bool in_tile = idx < nx * TY;
    const int x = in_tile ? idx % nx : 0, yl = in_tile ? idx / nx : 0;
    const int gy = y0 + yl, gw = wrap_y(gy);
    xq[q] = x;
    lv[q] = in_tile && gw >= 0;
    sv[q] = in_tile && gy < ny;
    oxy[q] = x + nx * (gw < 0 ? 0 : gw);
    lcol[q] = (NF + yl) * W + NF + x;
  }
  // halo rows: index i = t + 512 h over 2 NF rows x nx
  int hxy[H], hcol[H], hx[H];
  bool hv_ok[H];
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int i = t + 512 * h;
    const bool ok = i < 2 * NF * nx;
    const int r = ok ? i / nx : 0, x = ok ? i % nx : 0;
    const int row = r < NF ? r : TY + r;
    const int gw = wrap_y(y0 - NF + row);
    hv_ok[h] = ok && gw >= 0;
    hxy[h] = x + nx * (gw < 0 ? 0 : gw);
    hcol[h] = ok ? row * W + NF + x : -1;
    hx[h] = x;
  }
  // pad columns are zero for a non-periodic x axis and never rewritten
  if (!p.px)
    for (int i = t; i < (TY + 2 * NF) * 2 * NF; i += 512) {
      const int row = i / (2 * NF), c = i % (2 * NF);
      lds[row * W + (c < NF ? c : nx + c)] = 0.0;
    }
  auto put = [&](int col, int x, double v) {  // LDS write with the x-wrap copies
    lds[col] = v;
    if (p.px) {
      if (x < NF) lds[col + nx] = v;
      if (x >= nx - NF) lds[col - nx] = v;
    }
  };
  double col[P][2 * NF + 1];
  // MODE 6: xn, un of the next centre plane (a delay line one plane deep) and of the plane
  // being prefetched
  constexpr int UP = UPD ? P : 1;
  double xd[UP], ud[UP], xnx[UP], unx[UP];
#pragma unroll
  for (int m = 0; m < 2 * NF; m++) {
    const long zo = zoff(z0 - NF + m);
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (UPD && m == NF) {  // the first centre plane z0 (in range: zo >= 0)
        col[q][m + 1] = lv[q] ? ld3(zo + oxy[q], xd[q], ud[q]) : 0.0;
        if (!lv[q]) xd[q] = ud[q] = 0.0;
      } else {
        col[q][m + 1] = (lv[q] && zo >= 0) ? ld(zo + oxy[q]) : 0.0;
      }
    }
  }
  // MODE 2: the second band and its H at the centre, prefetched like V_eff
  constexpr int YP = COMB ? P : 1;
  double yf[YP], hf[YP];
  double nxt[P], hvv[H], vf[P];
  auto prefetch = [&](int z) {
    const long zn = zoff(z + NF), zh = (long)z * plane;
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (UPD) {
        nxt[q] = (lv[q] && zn >= 0) ? ld3(zn + oxy[q], xnx[q], unx[q]) : 0.0;
        if (!(lv[q] && zn >= 0)) xnx[q] = unx[q] = 0.0;
      } else {
        nxt[q] = (lv[q] && zn >= 0) ? ld(zn + oxy[q]) : 0.0;
      }
      if (COMB && !SYM) yf[q] = sv[q] ? in2[zh + oxy[q]] : 0.0;
      if (COMB) hf[q] = sv[q] ? hin[zh + oxy[q]] : 0.0;
      vf[q] = (p.veff && sv[q]) ? p.veff[zh + oxy[q]] : 0.0;
    }
#pragma unroll
    for (int h = 0; h < H; h++) hvv[h] = hv_ok[h] ? ld(zh + hxy[h]) : 0.0;
  };
  prefetch(z0);
  double dsum = 0.0, dsum2 = 0.0, dsum3 = 0.0, amax = 0.0, cmx = 0.0;
  for (int z = z0; z < z1; z++) {
    double vcur[P], ycur[YP], hcur[YP], xcen[UP], ucen[UP];
#pragma unroll
    for (int q = 0; q < P; q++) {
#pragma unroll
      for (int m = 0; m < 2 * NF; m++) col[q][m] = col[q][m + 1];
      col[q][2 * NF] = nxt[q];
      if (UPD) {  // radius 1: the plane just prefetched is the next centre
        xcen[q] = xd[q];
        ucen[q] = ud[q];
        xd[q] = xnx[q];
        ud[q] = unx[q];
      }
      if (COMB && !SYM) ycur[q] = yf[q];
      if (COMB) hcur[q] = hf[q];
      vcur[q] = vf[q];
      if (t + 512 * q < nx * TY) put(lcol[q], xq[q], col[q][NF]);
    }
#pragma unroll
    for (int h = 0; h < H; h++)
      if (hcol[h] >= 0) put(hcol[h], hx[h], hvv[h]);
    if (z + 1 < z1) prefetch(z + 1);
    __syncthreads();
    const long zh = (long)z * plane;
#pragma unroll
    for (int q = 0; q < P; q++) {
      if (!sv[q]) continue;
      const double c = col[q][NF];
      double s = (p.c0 + vcur[q]) * c;
      const double *lr = lds + lcol[q];
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cx[m] * (lr[-m] + lr[m]);
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cy[m] * (lr[-m * W] + lr[m * W]);
#pragma unroll
      for (int m = 1; m <= NF; m++) s += p.cz[m] * (col[q][NF - m] + col[q][NF + m]);
      if (COMB) s = fa * s + fb * hcur[q];                 // H (fa x + fb y)
      if (!MAXES && p.dotp) dsum += (SYM ? fa * c : COMB ? fa * c + fb * ycur[q] : c) * s;
      if (COMB) dsum2 += SYM ? fa * c * hcur[q] : ycur[q] * s;
      if (MAXES) cmx = fmax(cmx, fabs(c));
      if (NORMS) dsum3 += c * c;
      if (UPD) {
        dsum2 += xcen[q] * xcen[q];
        dsum3 += xcen[q] * ucen[q];
      }
      if (!NORMS && p.omax) amax = fmax(amax, fabs(s));
      const long r = zh + oxy[q];
      out[r] = p.accumulate ? out[r] + s : s;
      if (UPD) {
        p.out_x[(long)b * p.ld_out + r] = xcen[q];
        p.out_h[(long)b * p.ld_out + r] = ucen[q];
      }
    }
    __syncthreads();
  }
  if (!MAXES && p.dotp) {
    const double tot = block_sum<512>(dsum, red);
    if (t == 0) p.dotp[(long)b * (p.dot_ld ? p.dot_ld : per) + tile + (long)p.ntile * zch] = tot;
  }
  if (COMB || UPD) {
    const double tot = block_sum<512>(dsum2, red);
    if (t == 0) p.dotp2[(long)b * (p.dot_ld ? p.dot_ld : per) + tile + (long)p.ntile * zch] = tot;
  }
  if (NORMS || UPD) {
    const double tot = block_sum<512>(dsum3, red);
    if (t == 0) p.dotp3[(long)b * (p.dot_ld ? p.dot_ld : per) + tile + (long)p.ntile * zch] = tot;
  }
  if (MAXES) {
    const double m = block_max<512>(cmx, red);
    if (t == 0) atomicMax(p.cmax + b, (unsigned long long)__double_as_longlong(m));
  }
  if (!NORMS && p.omax) {  // non-negative doubles order like their bit patterns
    const double m = block_max<512>(amax, red);
    if (t == 0) atomicMax(p.omax + b, (unsigned long long)__double_as_longlong(m));
  }
}

// waves per SIMD the register budget is sized for: 4 where the z-columns fit in 128 VGPRs
// (a 512-thread block is 2 waves per SIMD, so only even counts matter)
template <int NF, int P, int MODE>
constexpr int stencil_waves() {
  if (NF > 4) return 2;
  if (MODE == 2 || MODE == 5) return P == 1 ? 4 : 2;
  return P <= 2 ? 4 : 2;
}

template <int NF, int P, int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(stencil_waves<NF, P, MODE>()))) void k_stencil_lds(SL p) {
  extern __shared__ double lds[];
  __shared__ double red[8];
  if (p.skip && *p.skip) return;  // uniform: a device-side loop has finished
  const int L = blockIdx.x, xcd = L & 7, j = L >> 3;
  const int per = p.ntile * p.nzch;
  int b, rem;
  if (p.nbands >= 8) {  // all tiles and z-ranges of a band on one XCD
    b = (j / per) * 8 + xcd;
    if (b >= p.nbands) return;  // uniform for the block
    rem = j % per;
  } else {  // fewer bands than XCDs: each XCD takes a contiguous eighth of the items
    const long items = (long)p.nbands * per, chunk = (items + 7) / 8, it = xcd * chunk + j;
    if (j >= chunk || it >= items) return;
    b = (int)(it / per);
    rem = (int)(it % per);
  }
  stencil_block<NF, P, MODE>(p, b, rem % p.ntile, rem / p.ntile, lds, red);
}

// Several independent grids in one persistent launch: work-group g takes items g, g + G,
// ... of the concatenated item lists (system-major; an item is one band x y-tile x z-range
// of one system, described by sl[s]).  start[s] = first item of system s, start[nsys] = total.
template <int NF, int P, int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(NF <= 4 && P <= 2 ? 4 : 2))) void k_stencil_batch(const SL *sl, const long *start, int nsys) {
  extern __shared__ double lds[];
  __shared__ double red[8];
  const long total = start[nsys];
  int s = 0;
  for (long it = blockIdx.x; it < total; it += gridDim.x) {
    while (s + 1 < nsys && start[s + 1] <= it) s++;  // items ascend, so s only moves forward
    const SL p = sl[s];
    if (p.skip && *p.skip) continue;
    const long j = it - start[s];
    const int per = p.ntile * p.nzch;
    const int b = (int)(j / per), rem = (int)(j % per);
    stencil_block<NF, P, MODE>(p, b, rem % p.ntile, rem / p.ntile, lds, red);
    __syncthreads();  // the next item reuses the LDS slab
  }
}

// ----------------------------------------------------------------------------- preconditioner
struct PP {
  int nx, ny, nz, px, py, pz, halo, nfin;
  long ex, ey;
  const double *in;
  long ld_in;
  double *out;
  long ld_out;
};
__device__ inline double pval(const PP &p, const double *base, int x, int y, int z) {
  SP s;
  s.nx = p.nx; s.ny = p.ny; s.nz = p.nz; s.px = p.px; s.py = p.py; s.pz = p.pz;
  s.halo = p.halo; s.nfin = p.nfin; s.ex = p.ex; s.ey = p.ey;
  return val(s, base, x, y, z);
}
// ksprecondition8f.F90:17-32, same term order.
__global__ __launch_bounds__(256) void k_precond(PP p) {
  const int nxb = (p.nx + 63) / 64;
  const int x = (blockIdx.x % nxb) * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.x / nxb, b = blockIdx.z;
  if (x >= p.nx || y >= p.ny) return;
  const double *v = p.in + (long)b * p.ld_in;
  const double w = 1.0 / 12.0;
  const double o = w * pval(p, v, x - 1, y, z) + w * pval(p, v, x, y - 1, z) +
                   w * pval(p, v, x, y, z - 1) + 0.5 * pval(p, v, x, y, z) +
                   w * pval(p, v, x, y, z + 1) + w * pval(p, v, x, y + 1, z) +
                   w * pval(p, v, x + 1, y, z);
  p.out[(long)b * p.ld_out + x + (long)p.nx * (y + (long)p.ny * z)] = o;
}

// Fused residual + preconditioner: r = sign*(h - eps*s) at the 7 points, out = K r.
struct PRP {
  int nx, ny, nz, px, py, pz;
  const double *h, *s;
  long ld;
  const double *eps;
  double sign;
  double *r, *out;
};
__device__ inline double rval(const PRP &p, const double *h, const double *s, double e, int x,
                              int y, int z) {
  if (x < 0 || x >= p.nx) { if (!p.px) return 0.0; x = x < 0 ? x + p.nx : x - p.nx; }
  if (y < 0 || y >= p.ny) { if (!p.py) return 0.0; y = y < 0 ? y + p.ny : y - p.ny; }
  if (z < 0 || z >= p.nz) { if (!p.pz) return 0.0; z = z < 0 ? z + p.nz : z - p.nz; }
  const long i = (long)x + (long)p.nx * ((long)y + (long)p.ny * z);
  return p.sign * (h[i] - e * s[i]);
}
__global__ __launch_bounds__(256) void k_precond_resid(PRP p) {
  const int nxb = (p.nx + 63) / 64;
  const int x = (blockIdx.x % nxb) * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.x / nxb, b = blockIdx.z;
  if (x >= p.nx || y >= p.ny) return;
  const double *h = p.h + (long)b * p.ld, *s = p.s + (long)b * p.ld;
  const double e = p.eps[b];
  const double w = 1.0 / 12.0;
  const double c = rval(p, h, s, e, x, y, z);
  const double o = w * rval(p, h, s, e, x - 1, y, z) + w * rval(p, h, s, e, x, y - 1, z) +
                   w * rval(p, h, s, e, x, y, z - 1) + 0.5 * c + w * rval(p, h, s, e, x, y, z + 1) +
                   w * rval(p, h, s, e, x, y + 1, z) + w * rval(p, h, s, e, x + 1, y, z);
  const long i = (long)b * p.ld + x + (long)p.nx * (y + (long)p.ny * z);
  if (p.r) p.r[i] = c;
  p.out[i] = o;
}

// ----------------------------------------------------------------------------- halo (leaf)
__global__ void k_halo_axis(double *v, int axis, int na, int nb, int nc, int nf, int nf1,
                            int nsnd) {
  const long ex = na + nf + nf1 + 1, ey = nb + nf + nf1 + 1;
  auto H = [&](int ia, int ib, int ic) -> long {
    return (long)(ia + nf1) + ex * ((long)(ib + nf1) + ey * (long)(ic + nf1));
  };
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  // (face index u in 1..d1, v in 1..d2, iov in 1..nsnd)
  long d1, d2;
  if (axis == 0) { d1 = nb; d2 = nc; } else if (axis == 1) { d1 = na; d2 = nc; } else { d1 = na; d2 = nb; }
  if (t >= d1 * d2 * nsnd) return;
  const int u = 1 + (int)(t % d1), w = 1 + (int)((t / d1) % d2), iov = 1 + (int)(t / (d1 * d2));
  if (axis == 0) {
    v[H(-nsnd + iov, u, w)] = v[H(na - nsnd + iov, u, w)];
    v[H(na + iov, u, w)] = v[H(iov, u, w)];
  } else if (axis == 1) {
    v[H(u, -nsnd + iov, w)] = v[H(u, nb - nsnd + iov, w)];
    v[H(u, nb + iov, w)] = v[H(u, iov, w)];
  } else {
    v[H(u, w, -nsnd + iov)] = v[H(u, w, nc - nsnd + iov)];
    v[H(u, w, nc + iov)] = v[H(u, w, iov)];
  }
}

// ----------------------------------------------------------------------------- nonlocal
// <p|psi> for active atom blockIdx.x and BB bands starting at blockIdx.y*BB.
template <int IPM, int BB>
__global__ __launch_bounds__(256) void k_gather(const int *act, int num_list, int nprjmx,
                                                const double *vnl, const int *lst,
                                                const double *psi, long ld, int nb,
                                                double *rp, long ld_rp, int accumulate, int band_major) {
  __shared__ double red[4][IPM * BB];
  // band_major: the band groups of one atom are consecutive work-groups, so the atom's
  // projector table is read from HBM once per XCD and then served by that XCD's L2 (2.60
  // against 2.81 ms per launch at 128^3 x 1024 bands x 128 atoms)
  const int a = band_major ? blockIdx.y : blockIdx.x;
  const int slot = act[5 * a], natinf = act[5 * a + 1], iprj = act[5 * a + 2];
  const int b0 = (band_major ? blockIdx.x : blockIdx.y) * BB;
  double acc[BB][IPM];
#pragma unroll
  for (int q = 0; q < BB; q++)
#pragma unroll
    for (int j = 0; j < IPM; j++) acc[q][j] = 0.0;
  const int *ls = lst + (long)num_list * slot;
  const double *pv = vnl + (long)num_list * nprjmx * slot;
  for (int k = threadIdx.x; k < natinf; k += 256) {
    const long g = ls[k] - 1;
    double w[BB];
#pragma unroll
    for (int q = 0; q < BB; q++) w[q] = (b0 + q < nb) ? psi[(long)(b0 + q) * ld + g] : 0.0;
#pragma unroll
    for (int j = 0; j < IPM; j++) {
      if (j < iprj) {
        const double v = pv[k + (long)num_list * j];
#pragma unroll
        for (int q = 0; q < BB; q++) acc[q][j] += v * w[q];
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < BB; q++)
#pragma unroll
    for (int j = 0; j < IPM; j++) {
      const double s = wave_sum(acc[q][j]);
      if (lane == 0) red[wv][q * IPM + j] = s;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < IPM * BB; t += 256) {
    const int q = t / IPM, j = t % IPM;
    if (j < iprj && b0 + q < nb) {
      const double s = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
      double *o = rp + (long)(b0 + q) * ld_rp + j + (long)nprjmx * slot;
      *o = accumulate ? *o + s : s;
    }
  }
}

// coef(j, slot, b) = dvinv * sum_i rp(i, slot, b) X(i, j)   (nonlocaloperation8f.F90:180-198)
// coef(j, slot, b) = dvinv * sum_i rp(i, slot, b) X(i, j); layout [b][slot][j] (ld_coef per
// band) or, with band_fastest, [slot][j][b] (the layout k_expand_rows reads).
__global__ void k_coef(const int *act, int nact, int nprjmx, int nums, int ns, const double *X,
                       int use_sss, const double *rp, long ld_rp, int nb, double dvinv,
                       double *coef, long ld_coef, int band_fastest) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long total = (long)nact * nprjmx * nb;
  if (t >= total) return;
  const int j = (int)(t % nprjmx), a = (int)((t / nprjmx) % nact), b = (int)(t / ((long)nprjmx * nact));
  const int slot = act[5 * a], iprj = act[5 * a + 2], na = act[5 * a + 3], ispe = act[5 * a + 4];
  if (j >= iprj) return;
  const double *Xa = use_sss ? X + (long)nprjmx * nprjmx * ispe
                             : X + (long)nprjmx * nprjmx * ((ns - 1) + (long)nums * na);
  const double *r = rp + (long)b * ld_rp + (long)nprjmx * slot;
  double s = 0.0;
  for (int i = 0; i < iprj; i++) s += r[i] * Xa[i + (long)nprjmx * j];
  const long jj = j + (long)nprjmx * slot;
  coef[band_fastest ? jj * nb + b : (long)b * ld_coef + jj] = s * dvinv;
}

__global__ void k_expand(const int *row_ptr, const int2 *ent, const int *iprj_slot,
                         const double *vnl, int num_list, int nprjmx, long n, const double *coef,
                         long ld_coef, int nb, double *out, long ld_out) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * nb) return;
  const long r = t % n;
  const int b = (int)(t / n);
  const int e0 = row_ptr[r], e1 = row_ptr[r + 1];
  if (e0 == e1) return;
  const double *cf = coef + (long)b * ld_coef;
  double s = 0.0;
  for (int e = e0; e < e1; e++) {
    const int2 en = ent[e];
    const int ip = iprj_slot[en.x];
    const double *v = vnl + en.y + (long)num_list * nprjmx * en.x;
    const double *cc = cf + (long)nprjmx * en.x;
    for (int j = 0; j < ip; j++) s += cc[j] * v[(long)num_list * j];
  }
  out[(long)b * ld_out + r] += s;
}

// Owner-computes expansion over the rows that carry projector entries, EB bands per
// thread: the entries' vnl values are read once per EB bands; per-band partials of
// <psi|added> support the fused Rayleigh quotient.
template <int EB>
__global__ __launch_bounds__(256) void k_expand_rows(const int *nzrows, long nnz, const int *row_ptr,
                                                     const int2 *ent, const int *iprj_slot,
                                                     const double *vnl, int num_list, int nprjmx,
                                                     const double *coef, int nb, double *out,
                                                     long ld_out, const double *psi, long ld_psi,
                                                     double *part, long part_ld, int part_off,
                                                     const double *in2, const double *ca,
                                                     const double *cb, double *part2,
                                                     double *maxpart, int band_major) {
  // coef is [slot][j][b] (bands fastest): the EB coefficients of one (slot, j) are contiguous
  // and shared by all rows of the same atom, so they come from L1.  band_major: the band
  // groups of one chunk of rows are consecutive work-groups, so the chunk's entries and
  // projector values are read from HBM about once per XCD instead of once per band group.
  __shared__ double red[EB][4];
  __shared__ double redm[EB][4];
  const unsigned rchunk = band_major ? blockIdx.y : blockIdx.x, nrchunk = band_major ? gridDim.y : gridDim.x;
  const long i = rchunk * 256L + threadIdx.x;
  const int b0 = (band_major ? blockIdx.x : blockIdx.y) * EB;
  const bool full = b0 + EB <= nb;
  double acc[EB];
#pragma unroll
  for (int q = 0; q < EB; q++) acc[q] = 0.0;
  long r = -1;
  if (i < nnz) {
    r = nzrows[i];
    const int e1 = row_ptr[r + 1];
    for (int e = row_ptr[r]; e < e1; e++) {
      const int2 en = ent[e];
      const int ip = iprj_slot[en.x];
      const double *v = vnl + en.y + (long)num_list * nprjmx * en.x;
      const double *cc = coef + (long)nprjmx * en.x * nb + b0;
      for (int j = 0; j < ip; j++) {
        const double vj = v[(long)num_list * j];
        const double *cj = cc + (long)j * nb;
        if (full) {
#pragma unroll
          for (int q = 0; q < EB; q++) acc[q] += cj[q] * vj;
        } else {
#pragma unroll
          for (int q = 0; q < EB; q++)
            if (b0 + q < nb) acc[q] += cj[q] * vj;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < EB; q++)
      if (b0 + q < nb) out[(long)(b0 + q) * ld_out + r] += acc[q];
  }
  if (maxpart) {  // per band: the work-group's largest |added value|
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < EB; q++) {
      double m = (r >= 0 && b0 + q < nb) ? fabs(acc[q]) : 0.0;
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
      if (lane == 0) redm[q][wv] = m;
    }
    __syncthreads();
    if (threadIdx.x < EB && b0 + (int)threadIdx.x < nb) {
      const int q = threadIdx.x;
      maxpart[(long)(b0 + q) * nrchunk + rchunk] =
          fmax(fmax(redm[q][0], redm[q][1]), fmax(redm[q][2], redm[q][3]));
    }
  }
  if (part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < EB; q++) {
      double d = 0.0;
      if (r >= 0 && b0 + q < nb) {
        const long o = (long)(b0 + q) * ld_psi + r;
        d = (in2 ? ca[b0 + q] * psi[o] + cb[b0 + q] * in2[o] : psi[o]) * acc[q];
      }
      d = wave_sum(d);
      if (lane == 0) red[q][wv] = d;
    }
    __syncthreads();
    if (threadIdx.x < EB && b0 + (int)threadIdx.x < nb) {
      const int q = threadIdx.x;
      part[(long)(b0 + q) * part_ld + part_off + rchunk] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
    }
    if (in2) {  // <in2|added>
      __syncthreads();
#pragma unroll
      for (int q = 0; q < EB; q++) {
        double d = (r >= 0 && b0 + q < nb) ? in2[(long)(b0 + q) * ld_psi + r] * acc[q] : 0.0;
        d = wave_sum(d);
        if (lane == 0) red[q][wv] = d;
      }
      __syncthreads();
      if (threadIdx.x < EB && b0 + (int)threadIdx.x < nb) {
        const int q = threadIdx.x;
        part2[(long)(b0 + q) * part_ld + part_off + rchunk] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
      }
    }
  }
}

// sum over the active (slot, j) of coef(slot, j, b) rp(slot, j, b), per band: 64 bands per
// work-group, the atoms split over its 4 waves and the 4 sums added in a fixed order.
__global__ __launch_bounds__(256) void k_proj_dots(const int *act, int nact, int nprjmx,
                                                   const double *coef, int nb, const double *rp,
                                                   long ld_rp, double *part, long part_ld,
                                                   int part_off) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (b < nb) {
    for (int a = w; a < nact; a += 4) {
      const int slot = act[5 * a], iprj = act[5 * a + 2];
      const long j0 = (long)nprjmx * slot;
      for (int j = 0; j < iprj; j++) acc += coef[(j0 + j) * nb + b] * rp[(long)b * ld_rp + j0 + j];
    }
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && b < nb)
    part[(long)b * part_ld + part_off] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// scale[b] = 2^(14 - e) with m_b < 2^e, m_b = smax_b (FP64 bits) + max over the expansion's
// partial maxima: a bound on max |stencil part + projector part| without a pass over the bands
__global__ __launch_bounds__(256) void k_bound_scale(const unsigned long long *smax,
                                                     const double *maxpart, int np, double *scale) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  double m = 0.0;
  if (maxpart)
    for (int i = threadIdx.x; i < np; i += 256) m = fmax(m, maxpart[(long)b * np + i]);
  m = block_max<256>(m, red);
  if (threadIdx.x == 0) {
    const double bnd = __longlong_as_double((long long)smax[b]) + m;
    double sc = 1.0;
    if (bnd > 0.0) {
      int e;
      frexp(bnd, &e);
      sc = ldexp(1.0, 14 - e);
    }
    scale[b] = sc;
  }
}

__global__ void k_flag_rows(const int *row_ptr, long n, int *flag) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r < n) flag[r] = row_ptr[r + 1] > row_ptr[r] ? 1 : 0;
}
__global__ void k_compact_rows(const int *flag, const int *pos, long n, int *rows) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r < n && flag[r]) rows[pos[r]] = (int)r;
}

// inverse map construction
// lstvec2 entries outside [1, n] raise *bad and are skipped (never dereferenced).
__global__ void k_csr_count(const int *act, int nact, const int *lst, int num_list, long n,
                            int *cnt, int *bad) {
  const int a = blockIdx.y;
  const int slot = act[5 * a], natinf = act[5 * a + 1];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < natinf; k += gridDim.x * blockDim.x) {
    const int r = lst[k + (long)num_list * slot] - 1;
    if (r < 0 || r >= n) {
      atomicOr(bad, 1);
      continue;
    }
    atomicAdd(&cnt[r], 1);
  }
}
__global__ void k_csr_fill(const int *act, const int *lst, int num_list, long n,
                           const int *row_ptr, int *cursor, int2 *ent) {
  const int a = blockIdx.y;
  const int slot = act[5 * a], natinf = act[5 * a + 1];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < natinf; k += gridDim.x * blockDim.x) {
    const int r = lst[k + (long)num_list * slot] - 1;
    if (r < 0 || r >= n) continue;
    const int pos = atomicAdd(&cursor[r], 1);
    ent[row_ptr[r] + pos] = make_int2(slot, k);
  }
}
__global__ void k_csr_sort(const int *row_ptr, int2 *ent, long n) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int e0 = row_ptr[r], e1 = row_ptr[r + 1];
  for (int i = e0 + 1; i < e1; i++) {
    const int2 v = ent[i];
    int j = i - 1;
    while (j >= e0 && (ent[j].x > v.x || (ent[j].x == v.x && ent[j].y > v.y))) {
      ent[j + 1] = ent[j];
      j--;
    }
    ent[j + 1] = v;
  }
}

// ----------------------------------------------------------------------------- reductions
__global__ __launch_bounds__(256) void k_dot_part(long n, const double *x, long ldx,
                                                  const double *y, long ldy, double *part,
                                                  int nparts) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  const long chunk = (n + nparts - 1) / nparts;
  const long i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  const double *xb = x + (long)b * ldx, *yb = y + (long)b * ldy;
  double s = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) s += xb[i] * yb[i];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[(long)b * nparts + blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_reduce_parts(const double *part, int nparts, double scale,
                                                      double *res) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[(long)b * nparts + i];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) res[b] = scale * s;
}

// ----------------------------------------------------------------------------- elementwise
// Band-wise elementwise kernels: grid (chunks, bands), each work-group strides over its band,
// so no per-element division by the band length.
// work-groups per band: about 8192 in all, at least 2048 elements per work-group
inline unsigned ew_chunks(long n, int nb) {
  const long want = (8192 + nb - 1) / std::max(1, nb);
  return (unsigned)std::max(1L, std::min(want, (n + 2047) / 2048));
}

__global__ __launch_bounds__(256) void k_axpby(long n, const double *a, double ca,
                                               const double *x, long ldx, const double *bb,
                                               double cb, const double *y, long ldy,
                                               double *out, long ldo) {
  const int b = blockIdx.y;
  const double fa = ca * (a ? a[b] : 1.0), fb = cb * (bb ? bb[b] : 1.0);
  const double *xb = x + (long)b * ldx, *yb = y ? y + (long)b * ldy : nullptr;
  double *ob = out + (long)b * ldo;
  const long stride = (long)gridDim.x * 256;
  if (yb) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) ob[i] = fa * xb[i] + fb * yb[i];
  } else {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) ob[i] = fa * xb[i];
  }
}
__global__ __launch_bounds__(256) void k_scale(long n, const double *f, double *x, long ldx) {
  const int b = blockIdx.y;
  const double fb = f[b];
  double *xb = x + (long)b * ldx;
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) xb[i] *= fb;
}
__global__ __launch_bounds__(256) void k_copy2d(long n, const double *x, long ldx, double *y,
                                                long ldy) {
  const int b = blockIdx.y;
  const double *xb = x + (long)b * ldx;
  double *yb = y + (long)b * ldy;
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) yb[i] = xb[i];
}

// P = x + t_b P (t null: P = x); part1 = sum s.P, part2 = sum P.P per work-group (new P).
__global__ __launch_bounds__(256) void k_pupdate_dots(long n, const double *x, long ld,
                                                      const double *t, double *P,
                                                      const double *sv, double *part1,
                                                      double *part2, int nparts) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  const long chunk = (n + nparts - 1) / nparts;
  const long i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  const double *xb = x + (long)b * ld, *sb = sv + (long)b * ld;
  double *pb = P + (long)b * ld;
  const double tb = t ? t[b] : 0.0;
  double s1 = 0.0, s2 = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double v = t ? xb[i] + tb * pb[i] : xb[i];
    pb[i] = v;
    s1 += sb[i] * v;
    s2 += v * v;
  }
  s1 = block_sum<256>(s1, red);
  s2 = block_sum<256>(s2, red);
  if (threadIdx.x == 0) {
    part1[(long)b * nparts + blockIdx.x] = s1;
    part2[(long)b * nparts + blockIdx.x] = s2;
  }
}

// x = a_b x + c_b y and u = a2_b u + c2_b w (a2, c2 null: a, c); part1 = sum x.x,
// part2 = sum x.u (new values).
__global__ __launch_bounds__(256) void k_pair_update_dots(long n, const double *a,
                                                          const double *c, double *x,
                                                          const double *y, double *u,
                                                          const double *w, long ld,
                                                          double *part1, double *part2,
                                                          int nparts, const double *a2,
                                                          const double *c2, double *xo,
                                                          double *uo) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  const long chunk = (n + nparts - 1) / nparts;
  const long i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  const double ab = a[b], cb = c[b];
  const double ab2 = a2 ? a2[b] : ab, cb2 = c2 ? c2[b] : cb;
  const double *xb = x + (long)b * ld, *ub = u + (long)b * ld;
  const double *yb = y + (long)b * ld, *wb = w + (long)b * ld;
  double *xob = xo + (long)b * ld, *uob = uo + (long)b * ld;
  double s1 = 0.0, s2 = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double xv = ab * xb[i] + cb * yb[i];
    const double uv = ab2 * ub[i] + cb2 * wb[i];
    xob[i] = xv;
    uob[i] = uv;
    s1 += xv * xv;
    s2 += xv * uv;
  }
  s1 = block_sum<256>(s1, red);
  s2 = block_sum<256>(s2, red);
  if (threadIdx.x == 0) {
    part1[(long)b * nparts + blockIdx.x] = s1;
    part2[(long)b * nparts + blockIdx.x] = s2;
  }
}
__global__ void k_fill(long n, double v, double *x) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}
__global__ void k_permute(long n, int nb, const int *perm, const double *x, long ldx, double *y,
                          long ldy) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * nb) return;
  const long i = t % n;
  const int b = (int)(t / n);
  y[(long)perm[b] * ldy + i] = x[(long)b * ldx + i];
}
// y(:, dst_j) = x(:, src_j) for column j of the lists (null list: identity)
__global__ __launch_bounds__(256) void k_copy_cols(long n, const int *src, const int *dst,
                                                   const double *x, long ldx, double *y, long ldy) {
  const int j = blockIdx.y;
  const double *xb = x + (long)(src ? src[j] : j) * ldx;
  double *yb = y + (long)(dst ? dst[j] : j) * ldy;
  const long stride = (long)gridDim.x * 256;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += stride) yb[i] = xb[i];
}
__global__ __launch_bounds__(256) void k_canon_sign(int n, double *z) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  double *c = z + (long)n * blockIdx.x;
  double best = -1.0;
  int idx = 0x7fffffff;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double a = fabs(c[i]);
    if (a > best || (a == best && i < idx)) { best = a; idx = i; }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = idx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = bv[threadIdx.x + s];
      const int oi = bi[threadIdx.x + s];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const bool flip = c[bi[0]] < 0.0;
  __syncthreads();
  if (flip)
    for (int i = threadIdx.x; i < n; i += 256) c[i] = -c[i];
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

// splitmix64 finaliser as a counter-based generator; Box-Muller on two 53-bit uniforms.
__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_fill_normal(double *x, long n, unsigned long long seed) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long a = mix64(seed * 0x100000001B3ull + 2ull * (unsigned long long)i);
  const unsigned long long b = mix64(seed * 0x100000001B3ull + 2ull * (unsigned long long)i + 1ull);
  const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  x[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

}  // namespace

void launch_fill_normal(double *x, long n, unsigned long long seed, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_normal, dim3(nblk(n, 256)), dim3(256), 0, s, x, n, seed);
  HIPCK(hipGetLastError());
}

// ============================================================================= launchers
static void fd_weights(int nf, double *a) {
  static const double t[9][9] = {
      {0},
      {-2.0, 1.0},
      {-5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0},
      {-49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0},
      {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0},
      {-36883.0 / 12600.0, 5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0},
      {-5369.0 / 1800.0, 12.0 / 7.0, -15.0 / 56.0, 10.0 / 189.0, -1.0 / 112.0, 2.0 / 1925.0,
       -1.0 / 16632.0},
      {-3.0235941043083900, 1.7500000000000000, -0.2916666666666667, 6.4814814814814815e-2,
       -1.3257575757575758e-2, 2.1212121212121212e-3, -2.2662522662522663e-4,
       1.1892869035726179e-5},
      {-3.0548441043083900, 1.7777777777777778, -0.3111111111111111, 7.5420875420875421e-2,
       -1.7676767676767677e-2, 3.4809634809634810e-3, -5.1800051800051800e-4,
       5.0742907885765029e-5, -2.4281274281274281e-6}};
  for (int m = 0; m < 9; m++) a[m] = (nf >= 1 && nf <= 8) ? t[nf][m] : 0.0;
}
void fd_coefficients(int nf, double *acfd) { fd_weights(nf, acfd); }

// Tile plan of the LDS kernel for grid g and nb bands (nx <= 256).
struct LdsPlan {
  int TY, ntile, nzch, zc, P;
};
static LdsPlan lds_plan(const Grid &g, int NF, int nb, int force_p = 0) {
  LdsPlan L;
  // P points per thread (param "stencil_p": 1, 2 or 4; default 2), TY rows of nx points
  int pp = force_p > 0 ? force_p : (int)param("stencil_p", 2.0);
  pp = pp >= 4 ? 4 : (pp >= 2 ? 2 : 1);
  if (NF > 4) pp = 1;  // the z-column of a wide stencil leaves registers for one point
  // few bands (a Poisson solve): one point per thread and short z-ranges, for parallelism
  const bool few = (long)nb * ((g.ny * g.nx + 512 * pp - 1) / (512 * pp)) < 1024;
  if (few && force_p == 0) pp = 1;
  L.TY = std::max(1, std::min(g.ny, 512 * pp / g.nx));
  L.ntile = (g.ny + L.TY - 1) / L.TY;
  L.P = (g.nx * L.TY + 511) / 512;
  if (L.P == 3) L.P = 4;
  // z-ranges: enough blocks to fill the chip, each range at least 4 NF planes (2 NF planes
  // for few bands: the extra warm-up planes are re-read from cache)
  const long blocks = (long)nb * L.ntile;
  const int minz = few ? 2 * NF : 4 * NF;
  int nz_ch = 1;
  while (blocks * nz_ch < 2048 && g.nz / (nz_ch * 2) >= minz) nz_ch *= 2;
  L.zc = (g.nz + nz_ch - 1) / nz_ch;
  L.nzch = (g.nz + L.zc - 1) / L.zc;
  return L;
}
static bool lds_ok(const Grid &g) { return g.nx <= 256 && g.nx >= 8 && g.ny >= 8 && g.nz >= 8; }

// points per thread of the combination-form stencil (param comb_p, default 1: measured faster
// than 2 points at two waves per SIMD)
static int comb_points() { return (int)param("comb_p", 1.0); }

int stencil_parts_lds(const Grid &g, int NF, int nb, bool comb = false) {
  LdsPlan L = lds_plan(g, NF, nb, comb ? comb_points() : 0);
  return L.ntile * L.nzch;
}

// points per thread of the fused band update (param upd_p, default 1: 23.2 ms per launch at
// 128^3 x 1024 bands against 27.1 ms with 2 and 38.9 ms with 4)
static int upd_points() { return (int)param("upd_p", 1.0); }

int update_precond_nparts(const Grid &g, int nb) {
  LdsPlan L = lds_plan(g, 1, nb, upd_points());
  return L.ntile * L.nzch;
}

template <int NF, int MODE>
static void launch_lds_nf(const SL &p, const LdsPlan &L, int nb, hipStream_t s) {
  const size_t shm = sizeof(double) * (size_t)(L.TY + 2 * NF) * (p.nx + 2 * NF);
  const long per = (long)L.ntile * L.nzch;
  // nb >= 8: band groups of 8 (one band per XCD); fewer: the items split over the 8 XCDs
  const unsigned grid = (unsigned)(nb >= 8 ? (long)(nb + 7) / 8 * 8 * per : 8 * ((nb * per + 7) / 8));
  switch (L.P) {
    case 1: hipLaunchKernelGGL((k_stencil_lds<NF, 1, MODE>), dim3(grid), dim3(512), shm, s, p); break;
    case 2: hipLaunchKernelGGL((k_stencil_lds<NF, 2, MODE>), dim3(grid), dim3(512), shm, s, p); break;
    default: hipLaunchKernelGGL((k_stencil_lds<NF, 4, MODE>), dim3(grid), dim3(512), shm, s, p); break;
  }
}

// Generic LDS launch: coefficients in p; in2/eps for the residual mode.
static void launch_lds(SL &p, const Grid &g, int NF, int nb, bool res, hipStream_t s) {
  // the combination form prefetches two more values per point (the second band and its H):
  // one point per thread keeps it within the register budget of four waves per SIMD
  LdsPlan L = lds_plan(g, NF, nb, p.ca ? comb_points() : 0);
  if (L.P > 4) throw Error(1, "stencil tile does not fit the thread block");
  p.TY = L.TY; p.ntile = L.ntile; p.nzch = L.nzch; p.zc = L.zc; p.nbands = nb;
  if (res) {
    if (NF != 1) throw Error(1, "fused residual stencil supports radius 1 only");
    launch_lds_nf<1, 1>(p, L, nb, s);
  } else if (p.ca && p.in2) {
    switch (NF) {
#define CASE(N) case N: launch_lds_nf<N, 2>(p, L, nb, s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
    }
  } else if (p.ca) {  // symmetric form: y is not read
    switch (NF) {
#define CASE(N) case N: launch_lds_nf<N, 5>(p, L, nb, s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
    }
  } else if (p.dotp3) {  // plain with norm partials
    if (p.cmax || p.omax) throw Error(1, "stencil: norms and maxima in one launch");
    switch (NF) {
#define CASE(N) case N: launch_lds_nf<N, 3>(p, L, nb, s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
    }
  } else if (p.cmax) {  // plain with column maxima
    if (p.dotp) throw Error(1, "stencil: maxima and partials in one launch");
    switch (NF) {
#define CASE(N) case N: launch_lds_nf<N, 4>(p, L, nb, s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
    }
  } else {
    switch (NF) {
#define CASE(N) case N: launch_lds_nf<N, 0>(p, L, nb, s); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
    }
  }
  HIPCK(hipGetLastError());
}

bool stencil_band_fastest_coef(const Grid &g, bool halo_in) {
  return !halo_in && lds_ok(g) && param("stencil_lds", 1.0) >= 1.0;
}

int stencil_nparts(const Grid &g, int nf, int nb, bool comb) {
  if (lds_ok(g) && param("stencil_lds", 1.0) >= 1.0) return stencil_parts_lds(g, nf, nb, comb);
  const int zc = 8;
  return (int)(nblk(g.nx, 64) * nblk(g.ny, 4) * nblk(g.nz, zc));
}

void launch_stencil(const StencilCall &c, hipStream_t s) {
  const Grid &g = *c.g;
  const int nf = c.nfh_coeffs > 0 ? c.nfh_coeffs : g.nf;
  if (!c.halo_in && (g.nx < nf || g.ny < nf || g.nz < nf))
    throw Error(1, "grid smaller than the finite-difference radius");
  SP p{};
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
  p.px = g.nperi >= 1; p.py = g.nperi >= 2; p.pz = g.nperi >= 3;
  p.halo = c.halo_in ? 1 : 0;
  p.nfin = g.nf;
  p.ex = g.nx + 2 * g.nf;
  p.ey = g.ny + 2 * g.nf;
  p.in = c.in; p.ld_in = c.ld_in; p.out = c.out; p.ld_out = c.ld_out;
  p.veff = c.veff;
  double a[9];
  fd_weights(nf, a);
  for (int m = 0; m < 9; m++) {
    p.cx[m] = c.scale_lap * a[m] / g.dx / g.dx * (-0.5);
    p.cy[m] = c.scale_lap * a[m] / g.dy / g.dy * (-0.5);
    p.cz[m] = c.scale_lap * a[m] / g.dz / g.dz * (-0.5);
  }
  p.c0 = p.cx[0] + p.cy[0] + p.cz[0];
  p.accumulate = c.accumulate ? 1 : 0;
  p.skip = c.skip;
  if (c.pl && c.coef && c.pl->nent > 0) {
    if (c.coef_band_fastest != stencil_band_fastest_coef(g, c.halo_in))
      throw Error(1, "stencil: projector coefficients in the wrong layout");
    p.row_ptr = c.pl->row_ptr; p.ent = c.pl->ent; p.vnl = c.pl->vnl;
    p.num_list = c.pl->num_list; p.nprjmx = c.pl->nprjmx; p.iprj_slot = c.pl->iprj_slot;
    p.coef = c.coef; p.ld_coef = c.ld_coef;
  }
  if (!c.halo_in && lds_ok(g) && param("stencil_lds", 1.0) >= 1.0) {
    SL q{};
    q.nx = g.nx; q.ny = g.ny; q.nz = g.nz; q.px = p.px; q.py = p.py; q.pz = p.pz;
    q.in = c.in; q.ld_in = c.ld_in; q.out = c.out; q.ld_out = c.ld_out; q.veff = c.veff;
    for (int m = 0; m < 9; m++) { q.cx[m] = p.cx[m]; q.cy[m] = p.cy[m]; q.cz[m] = p.cz[m]; }
    q.c0 = p.c0; q.accumulate = p.accumulate; q.skip = p.skip;
    q.row_ptr = p.row_ptr; q.ent = p.ent; q.vnl = p.vnl; q.coef = p.coef; q.ld_coef = p.ld_coef;
    q.num_list = p.num_list; q.nprjmx = p.nprjmx; q.iprj_slot = p.iprj_slot;
    q.dotp = c.dot_partial;
    q.dot_ld = c.dot_ld;
    q.cmax = c.in_max;
    q.dotp3 = c.norm_partial;
    q.omax = c.out_max;
    if (c.ca) {
      if (!c.dot_partial || !c.dot_partial2 || !c.hin || c.accumulate)
        throw Error(1, "stencil: the combination form needs H in2, both partial buffers and no accumulate");
      q.in2 = c.in2; q.ca = c.ca; q.cb = c.cb; q.dotp2 = c.dot_partial2; q.hin = c.hin;
    }
    if (c.tag) {  // band arrays read and written: in (+ in2 + H in2) and out
      const double arrays = c.ca ? (c.in2 ? 4.0 : 3.0) : 2.0;
      Timed t(c.tag, 8.0 * ((double)c.nb * arrays * g.n() + (c.veff ? g.n() : 0)));
      launch_lds(q, g, nf, c.nb, false, s);
    } else {
      launch_lds(q, g, nf, c.nb, false, s);
    }
    if (c.pl && c.coef && c.pl->nent > 0) {  // nonlocal part: separate owner-computes pass
      const bool comb = c.ca != nullptr, sym = comb && !c.in2;
      const long pld = c.dot_ld ? c.dot_ld : stencil_parts_lds(g, nf, c.nb, comb);
      const int off = stencil_parts_lds(g, nf, c.nb, comb);
      if (c.rp && c.dot_partial && (!comb || sym || c.rp2)) {
        launch_expand_rows(*c.pl, c.coef, c.nb, c.out, c.ld_out, nullptr, 0, nullptr, 0, 0, s,
                           nullptr, nullptr, nullptr, nullptr, c.expand_max);
        launch_proj_dots(*c.pl, c.coef, c.nb, c.rp, c.ld_rp, c.dot_partial, pld, off, s);
        // symmetric form: <ca x|H y> already holds the projector part (H y is complete)
        if (comb && !sym) launch_proj_dots(*c.pl, c.coef, c.nb, c.rp2, c.ld_rp, c.dot_partial2, pld, off, s);
      } else {
        if (sym) throw Error(1, "stencil: the symmetric combination form needs projector-space dots");
        launch_expand_rows(*c.pl, c.coef, c.nb, c.out, c.ld_out, c.in, c.ld_in, c.dot_partial, pld,
                           off, s, comb ? c.in2 : nullptr, c.ca, c.cb, c.dot_partial2, c.expand_max);
      }
    }
    return;
  }
  if (c.ca) throw Error(1, "stencil: the combination form runs on the tiled path only");
  if (c.in_max || c.out_max || c.expand_max || c.norm_partial)
    throw Error(1, "stencil: norm and max outputs run on the tiled path only");
  p.zc = 8;
  p.nzch = (int)nblk(g.nz, p.zc);
  p.dotp = c.dot_partial;
  p.dot_ld = c.dot_ld;
  if (c.nb > 65535) throw Error(1, "too many bands in one launch");
  dim3 grid(nblk(g.nx, 64) * p.nzch, nblk(g.ny, 4), (unsigned)c.nb);
  dim3 blk(64, 4, 1);
  switch (nf) {
#define CASE(N) case N: hipLaunchKernelGGL(k_stencil<N>, grid, blk, 0, s, p); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: throw Error(1, "error in kslaplacian_r! This subroutine is for nf in 1-8.");
  }
  HIPCK(hipGetLastError());
}

// ----------------------------------------------------------------------------- batched stencil
StencilBatchPlan plan_stencil_batch(const std::vector<StencilCall> &calls, bool res, int P,
                                    const std::string &tag) {
  StencilBatchPlan b;
  b.res = res;
  b.P = P >= 4 ? 4 : (P >= 2 ? 2 : 1);
  b.nsys = (int)calls.size();
  if (b.nsys == 0) return b;
  std::vector<SL> sl(b.nsys);
  std::vector<long> start(b.nsys + 1, 0);
  size_t shm = 0;
  for (int i = 0; i < b.nsys; i++) {
    const StencilCall &c = calls[i];
    const Grid &g = *c.g;
    const int nf = res ? 1 : (c.nfh_coeffs > 0 ? c.nfh_coeffs : g.nf);
    if (i == 0) b.nf = nf;
    if (nf != b.nf) throw Error(1, "batched stencil: all systems need the same radius");
    if (c.halo_in || !lds_ok(g)) throw Error(1, "batched stencil: grid outside the tiled path");
    SL q{};
    q.nx = g.nx; q.ny = g.ny; q.nz = g.nz;
    q.px = g.nperi >= 1; q.py = g.nperi >= 2; q.pz = g.nperi >= 3;
    q.in = c.in; q.ld_in = c.ld_in; q.out = c.out; q.ld_out = c.ld_out; q.skip = c.skip;
    q.accumulate = c.accumulate ? 1 : 0;
    if (res) {
      q.in2 = c.in2; q.eps = c.eps; q.sign = c.sign;
      q.c0 = 0.5;
      q.cx[1] = q.cy[1] = q.cz[1] = 1.0 / 12.0;
    } else {
      q.veff = c.veff;
      double a[9];
      fd_weights(nf, a);
      for (int m = 0; m < 9; m++) {
        q.cx[m] = c.scale_lap * a[m] / g.dx / g.dx * (-0.5);
        q.cy[m] = c.scale_lap * a[m] / g.dy / g.dy * (-0.5);
        q.cz[m] = c.scale_lap * a[m] / g.dz / g.dz * (-0.5);
      }
      q.c0 = q.cx[0] + q.cy[0] + q.cz[0];
    }
    // whole z-columns unless the system has too few tiles to share the chip
    LdsPlan L = lds_plan(g, nf, c.nb, b.P);
    L.zc = g.nz;
    L.nzch = 1;
    while ((long)c.nb * L.ntile * L.nzch < 64 && g.nz / (L.nzch * 2) >= 2 * nf) {
      L.nzch *= 2;
      L.zc = (g.nz + L.nzch - 1) / L.nzch;
      L.nzch = (g.nz + L.zc - 1) / L.zc;
    }
    if (L.P > b.P) throw Error(1, "batched stencil: tile does not fit the thread block");
    q.TY = L.TY; q.ntile = L.ntile; q.nzch = L.nzch; q.zc = L.zc; q.nbands = c.nb;
    q.dotp = c.dot_partial;
    q.dot_ld = c.dot_ld ? c.dot_ld : (long)L.ntile * L.nzch;
    sl[i] = q;
    start[i + 1] = start[i] + (long)c.nb * L.ntile * L.nzch;
    shm = std::max(shm, sizeof(double) * (size_t)(L.TY + 2 * nf) * (g.nx + 2 * nf));
  }
  b.total = start[b.nsys];
  b.shm = shm;
  b.sl = scratch(tag + ".sl", sizeof(SL) * b.nsys);
  b.start = (long *)scratch(tag + ".start", sizeof(long) * (b.nsys + 1));
  HIPCK(hipMemcpy(b.sl, sl.data(), sizeof(SL) * b.nsys, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(b.start, start.data(), sizeof(long) * (b.nsys + 1), hipMemcpyHostToDevice));
  return b;
}

// parts per band of system i's stencil partial sums in a batch plan (tiles x z-ranges)
long stencil_batch_parts(const Grid &g, int nf, int nb, int P) {
  LdsPlan L = lds_plan(g, nf, nb, P);
  int nzch = 1, zc = g.nz;
  while ((long)nb * L.ntile * nzch < 64 && g.nz / (nzch * 2) >= 2 * nf) {
    nzch *= 2;
    zc = (g.nz + nzch - 1) / nzch;
    nzch = (g.nz + zc - 1) / zc;
  }
  return (long)L.ntile * nzch;
}

void launch_stencil_batch(const StencilBatchPlan &b, int grid, hipStream_t s) {
  if (b.nsys == 0 || b.total == 0) return;
  const unsigned G = (unsigned)std::max(1L, std::min<long>(grid, b.total));
  const SL *sl = (const SL *)b.sl;
  if (b.res) {
    switch (b.P) {
      case 1: hipLaunchKernelGGL((k_stencil_batch<1, 1, 1>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
      case 2: hipLaunchKernelGGL((k_stencil_batch<1, 2, 1>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
      default: hipLaunchKernelGGL((k_stencil_batch<1, 4, 1>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
    }
  } else {
    if (b.nf != 4) throw Error(1, "batched stencil: radius 4 only");
    switch (b.P) {
      case 1: hipLaunchKernelGGL((k_stencil_batch<4, 1, 0>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
      case 2: hipLaunchKernelGGL((k_stencil_batch<4, 2, 0>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
      default: hipLaunchKernelGGL((k_stencil_batch<4, 4, 0>), dim3(G), dim3(512), b.shm, s, sl, b.start, b.nsys); break;
    }
  }
  HIPCK(hipGetLastError());
}

void launch_halo_fill(double *v, int nperi, int na, int nb, int nc, int nf, int nf1, int nsnd,
                      hipStream_t s) {
  if (nsnd > nf1 + 1 || nsnd > nf || (nperi >= 1 && nsnd > na) || (nperi >= 2 && nsnd > nb) ||
      (nperi >= 3 && nsnd > nc))
    throw Error(1, "overlap_finitedifference_r: nsnd exceeds the halo or the grid");
  for (int axis = 0; axis < 3; axis++) {
    if (nperi < axis + 1) continue;
    long faces = axis == 0 ? (long)nb * nc : axis == 1 ? (long)na * nc : (long)na * nb;
    long tot = faces * nsnd;
    hipLaunchKernelGGL(k_halo_axis, dim3(nblk(tot, 256)), dim3(256), 0, s, v, axis, na, nb, nc,
                       nf, nf1, nsnd);
  }
  HIPCK(hipGetLastError());
}

void launch_precond(const Grid &g, const double *in, long ld_in, bool halo_in, int nf_halo,
                    double *out, long ld_out, int nb, hipStream_t s) {
  PP p{};
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
  p.px = g.nperi >= 1; p.py = g.nperi >= 2; p.pz = g.nperi >= 3;
  p.halo = halo_in; p.nfin = nf_halo;
  p.ex = g.nx + 2 * nf_halo; p.ey = g.ny + 2 * nf_halo;
  p.in = in; p.ld_in = ld_in; p.out = out; p.ld_out = ld_out;
  if (nb > 65535) throw Error(1, "too many bands in one launch");
  hipLaunchKernelGGL(k_precond, dim3(nblk(g.nx, 64) * g.nz, nblk(g.ny, 4), (unsigned)nb),
                     dim3(64, 4), 0, s, p);
  HIPCK(hipGetLastError());
}

void launch_precond_resid_dot(const Grid &g, const double *h, const double *sp, long ld,
                              const double *eps, double sign, double *out, int nb,
                              double *dot_partial, hipStream_t s, unsigned long long *out_max) {
  Timed t("precond_resid", 8.0 * (double)g.n() * nb * 3.0);
  SL q{};
  q.nx = g.nx; q.ny = g.ny; q.nz = g.nz;
  q.px = g.nperi >= 1; q.py = g.nperi >= 2; q.pz = g.nperi >= 3;
  q.in = h; q.in2 = sp; q.ld_in = ld; q.eps = eps; q.sign = sign; q.out = out; q.ld_out = ld;
  q.c0 = 0.5;
  q.cx[1] = q.cy[1] = q.cz[1] = 1.0 / 12.0;
  q.dotp = dot_partial;
  q.omax = out_max;
  if (out_max) launch_fill(nb, 0.0, (double *)out_max, s);  // all-zero bits
  launch_lds(q, g, 1, nb, true, s);
}

void launch_update_precond(const Grid &g, const double *x, const double *y, const double *h,
                           const double *w, long ld, const double *ca, const double *cb,
                           const double *ha, const double *hb, const double *eps, double sign,
                           double *x_out, double *h_out, double *out, int nb, double *part_r,
                           double *part_xx, double *part_xh, long part_ld,
                           unsigned long long *out_max, hipStream_t s) {
  if (!lds_ok(g) || param("stencil_lds", 1.0) < 1.0)
    throw Error(1, "fused band update: tiled stencil path only");
  if (x_out == x || h_out == h || out == x || out == h)
    throw Error(1, "fused band update: outputs must not alias the bands it reads");
  Timed t("update_precond", 8.0 * (double)g.n() * nb * 7.0);
  SL q{};
  q.nx = g.nx; q.ny = g.ny; q.nz = g.nz;
  q.px = g.nperi >= 1; q.py = g.nperi >= 2; q.pz = g.nperi >= 3;
  q.in = x; q.in2 = y; q.hin = h; q.hin2 = w; q.ld_in = ld;
  q.ca = ca; q.cb = cb; q.ha = ha; q.hb = hb; q.eps = eps; q.sign = sign;
  q.out = out; q.ld_out = ld; q.out_x = x_out; q.out_h = h_out;
  q.c0 = 0.5;
  q.cx[1] = q.cy[1] = q.cz[1] = 1.0 / 12.0;
  q.dotp = part_r; q.dotp2 = part_xx; q.dotp3 = part_xh;
  q.omax = out_max;
  if (out_max) launch_fill(nb, 0.0, (double *)out_max, s);  // all-zero bits
  LdsPlan L = lds_plan(g, 1, nb, upd_points());
  // each band's partials: ntile * nzch of them at stride part_ld
  if ((long)L.ntile * L.nzch > part_ld)
    throw Error(1, "fused band update: partial buffer stride below the work-group count");
  q.dot_ld = part_ld;
  q.TY = L.TY; q.ntile = L.ntile; q.nzch = L.nzch; q.zc = L.zc; q.nbands = nb;
  launch_lds_nf<1, 6>(q, L, nb, s);
  HIPCK(hipGetLastError());
}

void launch_precond_resid(const Grid &g, const double *h, const double *sp, long ld,
                          const double *eps, double sign, double *r_out, double *out, int nb,
                          hipStream_t s) {
  if (!r_out && lds_ok(g) && param("stencil_lds", 1.0) >= 1.0) {
    launch_precond_resid_dot(g, h, sp, ld, eps, sign, out, nb, nullptr, s);
    return;
  }
  Timed t("precond_resid", 8.0 * (double)g.n() * nb * (r_out ? 4.0 : 3.0));
  PRP p{};
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
  p.px = g.nperi >= 1; p.py = g.nperi >= 2; p.pz = g.nperi >= 3;
  p.h = h; p.s = sp; p.ld = ld; p.eps = eps; p.sign = sign; p.r = r_out; p.out = out;
  if (nb > 65535) throw Error(1, "too many bands in one launch");
  hipLaunchKernelGGL(k_precond_resid, dim3(nblk(g.nx, 64) * g.nz, nblk(g.ny, 4), (unsigned)nb),
                     dim3(64, 4), 0, s, p);
  HIPCK(hipGetLastError());
}

void launch_gather(const NonlocalPlan &pl, const double *psi, long ld, int nb, double *rp,
                   long ld_rp, bool accumulate, hipStream_t s) {
  if (pl.nact == 0 || nb == 0) return;
  Timed t("nl_gather", 8.0 * ((double)nb * pl.nent + (double)pl.nprjmx * pl.nent));
  const bool bm = param("nl_gather_band_major", 1.0) >= 1.0 && pl.nact <= 65535;
  auto grid = [&](unsigned groups) { return bm ? dim3(groups, pl.nact) : dim3(pl.nact, groups); };
  if (pl.max_iprj <= 9) {
    hipLaunchKernelGGL((k_gather<9, 4>), grid(nblk(nb, 4)), dim3(256), 0, s, pl.act,
                       pl.num_list, pl.nprjmx, pl.vnl, pl.lst, psi, ld, nb, rp, ld_rp,
                       accumulate ? 1 : 0, bm ? 1 : 0);
  } else if (pl.max_iprj <= 32) {
    hipLaunchKernelGGL((k_gather<32, 1>), grid((unsigned)nb), dim3(256), 0, s, pl.act,
                       pl.num_list, pl.nprjmx, pl.vnl, pl.lst, psi, ld, nb, rp, ld_rp,
                       accumulate ? 1 : 0, bm ? 1 : 0);
  } else {
    throw Error(1, "nonlocal projectors: more than 32 per atom are not supported");
  }
  HIPCK(hipGetLastError());
}

void launch_coef(const NonlocalPlan &pl, const double *rp, long ld_rp, int nb, const double *X,
                 bool use_sss, int ns, double dvinv, double *coef, long ld_coef, hipStream_t s,
                 bool band_fastest) {
  if (pl.nact == 0 || nb == 0) return;
  const long tot = (long)pl.nact * pl.nprjmx * nb;
  hipLaunchKernelGGL(k_coef, dim3(nblk(tot, 256)), dim3(256), 0, s, pl.act, pl.nact, pl.nprjmx,
                     pl.nums, ns, X, use_sss ? 1 : 0, rp, ld_rp, nb, dvinv, coef, ld_coef,
                     band_fastest ? 1 : 0);
  HIPCK(hipGetLastError());
}

void launch_expand(const NonlocalPlan &pl, long n, const double *coef, long ld_coef, int nb,
                   double *out, long ld_out, hipStream_t s) {
  if (pl.nent == 0 || nb == 0) return;
  hipLaunchKernelGGL(k_expand, dim3(nblk(n * nb, 256)), dim3(256), 0, s, pl.row_ptr, pl.ent,
                     pl.iprj_slot, pl.vnl, pl.num_list, pl.nprjmx, n, coef, ld_coef, nb, out,
                     ld_out);
  HIPCK(hipGetLastError());
}

int expand_nparts(const NonlocalPlan &pl) { return (int)std::max<long>(1, (pl.nnz + 255) / 256); }

void launch_proj_dots(const NonlocalPlan &pl, const double *coef, int nb, const double *rp,
                      long ld_rp, double *part, long part_ld, int part_off, hipStream_t s) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_proj_dots, dim3(nblk(nb, 64)), dim3(256), 0, s, pl.act, pl.nact,
                     pl.nprjmx, coef, nb, rp, ld_rp, part, part_ld, part_off);
  HIPCK(hipGetLastError());
}

void launch_expand_rows(const NonlocalPlan &pl, const double *coef, int nb, double *out,
                        long ld_out, const double *psi, long ld_psi, double *part, long part_ld,
                        int part_off, hipStream_t s, const double *in2, const double *ca,
                        const double *cb, double *part2, double *maxpart) {
  if (nb == 0) return;
  const int np = expand_nparts(pl);
  if (pl.nnz == 0) {
    if (maxpart) launch_fill((long)nb * np, 0.0, maxpart, s);
    if (part)
      for (int b = 0; b < nb; b++) launch_fill(1, 0.0, part + (long)b * part_ld + part_off, s);
    if (part2)
      for (int b = 0; b < nb; b++) launch_fill(1, 0.0, part2 + (long)b * part_ld + part_off, s);
    return;
  }
  Timed t("nl_expand", 8.0 * (2.0 * (double)pl.nnz * nb + (double)pl.nprjmx * pl.nent));
  constexpr int EB = 16;
  // band groups fastest (param "nl_expand_band_major"; measured 4.82 against 4.65 ms per launch
  // at 128^3 x 1024 bands, so off by default) unless the row chunks exceed grid.y's range
  const bool bm = param("nl_expand_band_major", 0.0) >= 1.0 && np <= 65535;
  const dim3 grid = bm ? dim3(nblk(nb, EB), np) : dim3(np, nblk(nb, EB));
  hipLaunchKernelGGL(k_expand_rows<EB>, grid, dim3(256), 0, s, pl.nzrows, pl.nnz,
                     pl.row_ptr, pl.ent, pl.iprj_slot, pl.vnl, pl.num_list, pl.nprjmx, coef, nb,
                     out, ld_out, psi, ld_psi, part, part_ld, part_off, in2, ca, cb, part2, maxpart,
                     bm ? 1 : 0);
  HIPCK(hipGetLastError());
}

void launch_bound_scale(int nb, const unsigned long long *smax, const double *maxpart, int np,
                        double *scale, hipStream_t s) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_bound_scale, dim3(nb), dim3(256), 0, s, smax, maxpart, np, scale);
  HIPCK(hipGetLastError());
}

int dot_nparts(long n) {
  long p = (n + 8191) / 8192;
  if (p < 1) p = 1;
  if (p > 1024) p = 1024;
  return (int)p;
}

void launch_dots(long n, const double *x, long ldx, const double *y, long ldy, int nb,
                 double scale, double *res, hipStream_t s, double *part) {
  if (nb == 0) return;
  const int np = dot_nparts(n);
  Timed t("dots", (x == y && ldx == ldy ? 8.0 : 16.0) * (double)n * nb);
  if (!part) part = (double *)scratch("dots.part", sizeof(double) * (size_t)np * nb);
  hipLaunchKernelGGL(k_dot_part, dim3(np, nb), dim3(256), 0, s, n, x, ldx, y, ldy, part, np);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part, np, scale, res);
  HIPCK(hipGetLastError());
}

void launch_reduce_partials(const double *part, int nparts, int nb, double scale, double *res,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part, nparts, scale, res);
  HIPCK(hipGetLastError());
}

void launch_axpby(long n, int nb, const double *a, double ca, const double *x, long ldx,
                  const double *b, double cb, const double *y, long ldy, double *out, long ldo,
                  hipStream_t s) {
  if (n * nb == 0) return;
  Timed t("axpby", 8.0 * (double)n * nb * (y ? 3.0 : 2.0));
  hipLaunchKernelGGL(k_axpby, dim3(ew_chunks(n, nb), nb), dim3(256), 0, s, n, a, ca, x, ldx, b,
                     cb, y, ldy, out, ldo);
  HIPCK(hipGetLastError());
}
void launch_scale(long n, int nb, const double *f, double *x, long ldx, hipStream_t s) {
  if (n * nb == 0) return;
  hipLaunchKernelGGL(k_scale, dim3(ew_chunks(n, nb), nb), dim3(256), 0, s, n, f, x, ldx);
  HIPCK(hipGetLastError());
}
void launch_copy2d(long n, int nb, const double *x, long ldx, double *y, long ldy, hipStream_t s) {
  if (n * nb == 0) return;
  hipLaunchKernelGGL(k_copy2d, dim3(ew_chunks(n, nb), nb), dim3(256), 0, s, n, x, ldx, y, ldy);
  HIPCK(hipGetLastError());
}
void launch_pupdate_dots(long n, int nb, const double *x, const double *t, double *P, long ld,
                         const double *sv, double scale, double *res1, double *res2,
                         hipStream_t s) {
  if (n * nb == 0) return;
  const int np = dot_nparts(n);
  Timed tt("p_update_dots", 8.0 * (double)n * nb * (t ? 4.0 : 3.0));
  double *part = (double *)scratch("pud.part", sizeof(double) * (size_t)2 * np * nb);
  hipLaunchKernelGGL(k_pupdate_dots, dim3(np, nb), dim3(256), 0, s, n, x, ld, t, P, sv, part,
                     part + (long)np * nb, np);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part, np, scale, res1);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part + (long)np * nb, np, scale,
                     res2);
  HIPCK(hipGetLastError());
}
void launch_pair_update_dots(long n, int nb, const double *a, const double *c, double *x,
                             const double *y, double *u, const double *w, long ld, double scale,
                             double *res1, double *res2, hipStream_t s, const double *a2,
                             const double *c2, double *x_out, double *u_out) {
  if (n * nb == 0) return;
  if (!x_out) x_out = x;
  if (!u_out) u_out = u;
  const int np = dot_nparts(n);
  Timed tt("band_update_dots", 8.0 * (double)n * nb * 6.0);
  double *part = (double *)scratch("pair.part", sizeof(double) * (size_t)2 * np * nb);
  hipLaunchKernelGGL(k_pair_update_dots, dim3(np, nb), dim3(256), 0, s, n, a, c, x, y, u, w, ld,
                     part, part + (long)np * nb, np, a2, c2, x_out, u_out);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part, np, scale, res1);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, s, part + (long)np * nb, np, scale,
                     res2);
  HIPCK(hipGetLastError());
}
void launch_fill(long n, double v, double *x, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, s, n, v, x);
  HIPCK(hipGetLastError());
}
void launch_permute_cols(long n, int nb, const int *perm, const double *x, long ldx, double *y,
                         long ldy, hipStream_t s) {
  if (n * nb == 0) return;
  hipLaunchKernelGGL(k_permute, dim3(nblk(n * nb, 256)), dim3(256), 0, s, n, nb, perm, x, ldx, y,
                     ldy);
  HIPCK(hipGetLastError());
}
void launch_copy_cols(long n, int cnt, const int *src, const int *dst, const double *x, long ldx,
                      double *y, long ldy, hipStream_t s) {
  if (n * cnt == 0) return;
  hipLaunchKernelGGL(k_copy_cols, dim3(ew_chunks(n, cnt), cnt), dim3(256), 0, s, n, src, dst, x, ldx,
                     y, ldy);
  HIPCK(hipGetLastError());
}
void launch_canon_sign(int n, double *z, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_canon_sign, dim3(n), dim3(256), 0, s, n, z);
  HIPCK(hipGetLastError());
}

// ----------------------------------------------------------------------------- plan build
namespace {
// order-dependent checksum of an int array (wrap-around integer sums: the same for any schedule)
__global__ void k_int_checksum(const int *a, long n, unsigned long long *out) {
  unsigned long long h = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    h += (unsigned long long)(unsigned)a[i] * (2654435761ull * (unsigned long long)(i + 1) + 97531ull);
  for (int o = 32; o > 0; o >>= 1) h += __shfl_down(h, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

// a built plan and what it was built from: the atom tables, the sizes, lstvec2's address and
// checksum, and the scratch buffers it lives in (a reallocated buffer invalidates it)
struct CachedPlan {
  std::vector<long> key;
  const int *lst = nullptr;
  unsigned long long sum = 0;
  std::vector<void *> bufs;
  NonlocalPlan pl;
};
std::map<std::string, CachedPlan> &plan_cache() {
  static std::map<std::string, CachedPlan> m;
  return m;
}
}  // namespace

void build_plan(NonlocalPlan &pl, const Grid &g, int natom, int nprjmx, int num_spe,
                int num_ppcell, int num_list, int nums, int key_in, int key_inps,
                const int *nprj, const int *indspe, const int *natinf, const int *natpri,
                const int *naps, const double *vnlocp_dev, const int *lstvec2_dev,
                const std::string &tag) {
  hipStream_t s = lib().stream;
  pl.natom = natom; pl.nprjmx = nprjmx; pl.num_list = num_list; pl.num_ppcell = num_ppcell;
  pl.nums = nums; pl.num_spe = num_spe; pl.vnl = vnlocp_dev; pl.lst = lstvec2_dev;
  std::vector<int> h_nprj = host_ints(nprj, num_spe), h_ind = host_ints(indspe, natom),
                   h_inf = host_ints(natinf, natom), h_pri = host_ints(natpri, natom),
                   h_aps = host_ints(naps, natom);
  // the plan of the previous call with the same inputs is reused (param "plan_cache", default
  // on): same sizes and atom tables, same lstvec2 address and contents
  const bool use_cache = param("plan_cache", 1.0) >= 1.0 && !lib().capturing;
  std::vector<long> key = {g.n(), g.nx, g.ny, g.nz, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_in,
                           key_inps};
  for (const std::vector<int> *v : {&h_nprj, &h_ind, &h_inf, &h_pri, &h_aps}) key.insert(key.end(), v->begin(), v->end());
  unsigned long long sum = 0;
  if (use_cache) {
    unsigned long long *d = (unsigned long long *)scratch(tag + ".sum", sizeof(unsigned long long));
    HIPCK(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    const long nl = (long)num_list * num_ppcell;
    hipLaunchKernelGGL(k_int_checksum, dim3((unsigned)std::min<long>(1024, std::max<long>(1, nblk(nl, 256)))),
                       dim3(256), 0, s, lstvec2_dev, nl, d);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(&sum, d, sizeof(sum), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    auto it = plan_cache().find(tag);
    if (it != plan_cache().end() && it->second.key == key && it->second.lst == lstvec2_dev && it->second.sum == sum) {
      const CachedPlan &c = it->second;
      const NonlocalPlan &q = c.pl;
      const long n = g.n();
      const std::vector<void *> now = {
          scratch(tag + ".act", sizeof(int) * std::max<size_t>(1, q.act_h.size())),
          scratch(tag + ".iprj", sizeof(int) * std::max(1, num_ppcell)),
          scratch(tag + ".rowptr", sizeof(int) * (n + 1)),
          scratch(tag + ".ent", sizeof(int2) * std::max<long>(1, q.nent)),
          scratch(tag + ".nzrows", sizeof(int) * std::max<long>(1, q.nnz))};
      if (now == c.bufs) {
        pl = q;
        pl.vnl = vnlocp_dev;
        pl.lst = lstvec2_dev;
        return;
      }
    }
  }
  pl.act_h.clear();
  std::vector<int> iprj_slot(num_ppcell > 0 ? num_ppcell : 1, 0);
  pl.max_iprj = 0;
  pl.max_natinf = 0;
  for (int na = 0; na < natom; na++) {
    if (h_pri[na] != key_in && h_pri[na] != key_inps) continue;
    const int slot = h_aps[na] - 1, ispe = h_ind[na] - 1;
    const int ip = h_nprj[ispe];
    if (slot < 0 || slot >= num_ppcell) throw Error(1, "naps out of range");
    if (h_inf[na] > num_list) throw Error(1, "error! num_list is too small.");
    pl.act_h.insert(pl.act_h.end(), {slot, h_inf[na], ip, na, ispe});
    iprj_slot[slot] = ip;
    pl.max_iprj = std::max(pl.max_iprj, ip);
    pl.max_natinf = std::max(pl.max_natinf, h_inf[na]);
  }
  pl.nact = (int)(pl.act_h.size() / 5);
  pl.act = (int *)scratch(tag + ".act", sizeof(int) * std::max<size_t>(1, pl.act_h.size()));
  pl.iprj_slot = (int *)scratch(tag + ".iprj", sizeof(int) * iprj_slot.size());
  if (!pl.act_h.empty())
    HIPCK(hipMemcpyAsync(pl.act, pl.act_h.data(), sizeof(int) * pl.act_h.size(),
                         hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(pl.iprj_slot, iprj_slot.data(), sizeof(int) * iprj_slot.size(),
                       hipMemcpyHostToDevice, s));
  const long n = g.n();
  long nent = 0;
  for (int a = 0; a < pl.nact; a++) nent += pl.act_h[5 * a + 1];
  pl.nent = nent;
  pl.row_ptr = (int *)scratch(tag + ".rowptr", sizeof(int) * (n + 1));
  int *cnt = (int *)scratch(tag + ".cnt", sizeof(int) * (n + 1));
  pl.ent = (int2 *)scratch(tag + ".ent", sizeof(int2) * std::max<long>(1, nent));
  HIPCK(hipMemsetAsync(cnt, 0, sizeof(int) * (n + 1), s));
  if (pl.nact > 0) {
    int *bad = (int *)scratch(tag + ".bad", sizeof(int));
    HIPCK(hipMemsetAsync(bad, 0, sizeof(int), s));
    dim3 grid(nblk(pl.max_natinf, 256), pl.nact);
    hipLaunchKernelGGL(k_csr_count, grid, dim3(256), 0, s, pl.act, pl.nact, pl.lst, num_list, n,
                       cnt, bad);
    HIPCK(hipGetLastError());
    int hbad = 0;
    HIPCK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (hbad) throw Error(1, "lstvec2 holds a grid index outside 1..ncpx*ncpy*ncpz");
  }
  size_t tmp_bytes = 0;
  HIPCK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, pl.row_ptr, (int)(n + 1), s));
  void *tmp = scratch(tag + ".scan", tmp_bytes);
  HIPCK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, pl.row_ptr, (int)(n + 1), s));
  if (pl.nact > 0) {
    HIPCK(hipMemsetAsync(cnt, 0, sizeof(int) * (n + 1), s));
    dim3 grid(nblk(pl.max_natinf, 256), pl.nact);
    hipLaunchKernelGGL(k_csr_fill, grid, dim3(256), 0, s, pl.act, pl.lst, num_list, n, pl.row_ptr,
                       cnt, pl.ent);
    hipLaunchKernelGGL(k_csr_sort, dim3(nblk(n, 256)), dim3(256), 0, s, pl.row_ptr, pl.ent, n);
    HIPCK(hipGetLastError());
  }
  // compacted list of rows with entries
  int *flag = cnt;  // reuse (n+1 ints)
  int *pos = (int *)scratch(tag + ".pos", sizeof(int) * (n + 1));
  hipLaunchKernelGGL(k_flag_rows, dim3(nblk(n, 256)), dim3(256), 0, s, pl.row_ptr, n, flag);
  HIPCK(hipMemsetAsync(flag + n, 0, sizeof(int), s));
  HIPCK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flag, pos, (int)(n + 1), s));
  int hnnz = 0;
  HIPCK(hipMemcpyAsync(&hnnz, pos + n, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  pl.nnz = hnnz;
  pl.nzrows = (int *)scratch(tag + ".nzrows", sizeof(int) * std::max<long>(1, pl.nnz));
  hipLaunchKernelGGL(k_compact_rows, dim3(nblk(n, 256)), dim3(256), 0, s, flag, pos, n, pl.nzrows);
  HIPCK(hipGetLastError());
  if (use_cache) {
    CachedPlan &c = plan_cache()[tag];
    c.key = std::move(key);
    c.lst = lstvec2_dev;
    c.sum = sum;
    c.pl = pl;
    c.bufs = {pl.act, pl.iprj_slot, pl.row_ptr, pl.ent, pl.nzrows};
  }
}

}  // namespace kukan
