// bloch.hip — the complex (Bloch, k-point) path: nrc = 1, collinear spin (ncol = 1), no
// spin-orbit (nso = 0).  Bands are complex*16 (interleaved re, im), k = (skpx, skpy, skpz).
//
// H_k u = e^{-i k.r} H e^{i k.r} u for the cell-periodic part u.  The reference multiplies a
// halo'd copy of u by e^{i k.r} (unwrapped coordinates), applies the real stencil and multiplies
// back (kslaplacian8f.F90:560-657); on the interior of a periodic grid that is the "twisted"
// stencil sum_m c_m (e^{i k m h} u(x + m) + e^{-i k m h} u(x - m)), which the band kernels here
// apply directly with wrapped indices (one pass, no halo buffer, no per-point sincos).  The
// leaf entry point kukan_kslaplacian_c keeps the reference's halo'd form and its side effect
// (the phase left on the input).  Projectors carry the phase of their point's unwrapped
// position, e^{i k.x} with x = (lst + nat) h - h/2 (nonlocaloperation8f.F90:95-107): a phase
// table per call, the gather <p|u> = sum vnlocp e^{i k.x} u and the owner-computes expansion
// sum e^{-i k.x} vnlocp c over the same CSR plan as the real path.
//
// Solvers: kscg_c band by band as kscg8f.F90:492-1032 (device kernels, per-band scalars on
// the host), orthogonalization_c as Cholesky-QR (rocBLAS zgemm, rocSOLVER zpotrf, ztrsm: the
// Gram-Schmidt Q), single-band Gram-Schmidt with zgemv, rayleighritz_c with zgemm and
// rocSOLVER zheevd.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "kukan_hip.h"
#include "kukan_internal.hpp"
#include "solvers.hpp"
#include "bloch.hpp"

namespace kukan {
namespace {

typedef double2 zt;
typedef std::complex<double> zc;

inline unsigned nb_(long n, int t) { return (unsigned)((n + t - 1) / t); }

__host__ __device__ inline zt zmk(double r, double i) { return make_double2(r, i); }
__device__ inline zt zmul(zt a, zt b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline zt zadd(zt a, zt b) { return zmk(a.x + b.x, a.y + b.y); }
__device__ inline zt zsc(double s, zt a) { return zmk(s * a.x, s * a.y); }
__device__ inline zt zcj(zt a) { return zmk(a.x, -a.y); }

// ----------------------------------------------------------------------------- phases
// ph(i, slot) = e^{i k.x_i} for the list entries of every active atom
__global__ void k_zphase(const int *act, int nact, int num_list, const int *lstx, const int *lsty,
                         const int *lstz, const int *natx, const int *naty, const int *natz,
                         double dx, double dy, double dz, double kx, double ky, double kz, zt *ph) {
  const int a = blockIdx.y;
  if (a >= nact) return;
  const int slot = act[5 * a], natinf = act[5 * a + 1], na = act[5 * a + 3];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < natinf; i += gridDim.x * blockDim.x) {
    const size_t k = (size_t)i + (size_t)num_list * slot;
    const int ix = lstx[k] + natx[na], iy = lsty[k] + naty[na], iz = lstz[k] + natz[na];
    const double xx = ix * dx - 0.5 * dx, yy = iy * dy - 0.5 * dy, zz = iz * dz - 0.5 * dz;
    const double th = kx * xx + ky * yy + kz * zz;
    double s, c;
    sincos(th, &s, &c);
    ph[k] = zmk(c, s);
  }
}

// ----------------------------------------------------------------------------- gather
// cp(j, slot, b) (=|+=) sum_i vnlocp(i, j, slot) ph(i, slot) u_b(lst(i, slot)): one work-group
// per (active atom, band), partial sums per thread, a fixed-order tree reduction.
constexpr int ZG_T = 256, ZG_J = 16;
__global__ __launch_bounds__(ZG_T) void k_zgather(const int *act, int num_list, int nprjmx,
                                                  const double *vnl, const int *lst, const zt *ph,
                                                  const zt *u, long ldu, zt *cp, long ldcp,
                                                  int accumulate) {
  __shared__ zt red[ZG_T];
  const int a = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int slot = act[5 * a], natinf = act[5 * a + 1], iprj = act[5 * a + 2];
  const zt *ub = u + ldu * b;
  zt acc[ZG_J];
#pragma unroll
  for (int j = 0; j < ZG_J; j++) acc[j] = zmk(0.0, 0.0);
  for (int i = t; i < natinf; i += ZG_T) {
    const size_t k = (size_t)i + (size_t)num_list * slot;
    const zt v = zmul(ph[k], ub[lst[k] - 1]);
#pragma unroll
    for (int j = 0; j < ZG_J; j++)
      if (j < iprj) acc[j] = zadd(acc[j], zsc(vnl[(size_t)i + (size_t)num_list * (j + (size_t)nprjmx * slot)], v));
  }
  for (int j = 0; j < iprj; j++) {
    red[t] = acc[j < ZG_J ? j : 0];
    __syncthreads();
    for (int w = ZG_T / 2; w > 0; w >>= 1) {
      if (t < w) red[t] = zadd(red[t], red[t + w]);
      __syncthreads();
    }
    if (t == 0) {
      zt *o = cp + ldcp * b + (size_t)j + (size_t)nprjmx * slot;
      *o = accumulate ? zadd(*o, red[0]) : red[0];
    }
    __syncthreads();
  }
}

// coef(i, slot, b) = sum_j X(i, j) (cp(j, slot, b) / dV) in the reference's order
// (nonlocaloperation8f.F90:300-331: ctmpv = cpsep dxyzin, ctmpv2(i) += ctmpv(j) X(i,j)).
// X = dijsss(:, :, i3, na) (nhs = 1, i3 = ns1) or (:, :, 1, species) (nhs = 0).
__global__ void k_zcoef(const int *act, int nact, int nprjmx, int nhs, int d3, int i3,
                        const double *X, const zt *cp, long ldcp, double dxyzin, zt *coef,
                        long ldco) {
  const int a = blockIdx.x, b = blockIdx.y, i = threadIdx.x;
  if (a >= nact) return;
  const int slot = act[5 * a], iprj = act[5 * a + 2], na = act[5 * a + 3], ispe = act[5 * a + 4];
  if (i >= iprj) return;
  const int i4 = nhs == 1 ? na : ispe;
  const double *Xa = X + (size_t)nprjmx * nprjmx * ((size_t)(i3 - 1) + (size_t)d3 * i4);
  const zt *c = cp + ldcp * b + (size_t)nprjmx * slot;
  zt s = zmk(0.0, 0.0);
  for (int j = 0; j < iprj; j++) s = zadd(s, zsc(Xa[i + (size_t)nprjmx * j], zsc(dxyzin, c[j])));
  coef[ldco * b + (size_t)i + (size_t)nprjmx * slot] = s;
}

// out_b(r) += sum over (slot, k) of row r of conj(ph(k, slot)) sum_j coef(j, slot, b)
// vnlocp(k, j, slot); one thread per (touched row, band)
__global__ void k_zexpand(const int *nzrows, long nnz, const int *row_ptr, const int2 *ent,
                          int num_list, int nprjmx, const int *iprj_slot, const double *vnl,
                          const zt *ph, const zt *coef, long ldco, zt *out, long ldo) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (q >= nnz) return;
  const int r = nzrows[q];
  zt s = out[ldo * b + r];
  for (int e = row_ptr[r]; e < row_ptr[r + 1]; e++) {
    const int slot = ent[e].x, k = ent[e].y, iprj = iprj_slot[slot];
    const zt *c = coef + ldco * b + (size_t)nprjmx * slot;
    const double *p = vnl + (size_t)k + (size_t)num_list * (size_t)nprjmx * slot;
    zt av = zsc(p[0], c[0]);
    for (int j = 1; j < iprj; j++) av = zadd(av, zsc(p[(size_t)num_list * j], c[j]));
    s = zadd(s, zmul(zcj(ph[(size_t)k + (size_t)num_list * slot]), av));
  }
  out[ldo * b + r] = s;
}

// ----------------------------------------------------------------------------- stencils
struct ZLap {
  int nx, ny, nz, nf, px, py, pz;
  double c0;
  double cx[9], cy[9], cz[9];
  zt ex[9], ey[9], ez[9];  // e^{i k m h} per axis
};

// periodic or zero-padded neighbour along one axis
__device__ inline int nbr(int i, int m, int n, int per) {
  int j = i + m;
  if (j >= 0 && j < n) return j;
  if (!per) return -1;
  j %= n;
  return j < 0 ? j + n : j;
}

// out_b (=|+=) H_k u_b on the interior: sweeps x (with the centre and veff), y, z, each summed
// minus side (m = nf .. 1), [centre,] plus side (m = 1 .. nf) as kslaplacian_c_01
__global__ void k_zlap(ZLap L, const zt *u, long ldu, const double *veff, zt *out, long ldo,
                       int accumulate) {
  const long n = (long)L.nx * L.ny * L.nz;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= n) return;
  const int ix = (int)(p % L.nx), iy = (int)((p / L.nx) % L.ny), iz = (int)(p / ((long)L.nx * L.ny));
  const zt *ub = u + ldu * b;
  auto at = [&](int x, int y, int z) { return ub[x + (long)L.nx * (y + (long)L.ny * z)]; };
  zt s = accumulate ? out[ldo * b + p] : zmk(0.0, 0.0);
  for (int m = L.nf; m >= 1; m--) {
    const int j = nbr(ix, -m, L.nx, L.px);
    if (j >= 0) s = zadd(s, zsc(L.cx[m], zmul(zcj(L.ex[m]), at(j, iy, iz))));
  }
  s = zadd(s, zsc(L.c0 + (veff ? veff[p] : 0.0), at(ix, iy, iz)));
  for (int m = 1; m <= L.nf; m++) {
    const int j = nbr(ix, m, L.nx, L.px);
    if (j >= 0) s = zadd(s, zsc(L.cx[m], zmul(L.ex[m], at(j, iy, iz))));
  }
  for (int m = L.nf; m >= 1; m--) {
    const int j = nbr(iy, -m, L.ny, L.py);
    if (j >= 0) s = zadd(s, zsc(L.cy[m], zmul(zcj(L.ey[m]), at(ix, j, iz))));
  }
  for (int m = 1; m <= L.nf; m++) {
    const int j = nbr(iy, m, L.ny, L.py);
    if (j >= 0) s = zadd(s, zsc(L.cy[m], zmul(L.ey[m], at(ix, j, iz))));
  }
  for (int m = L.nf; m >= 1; m--) {
    const int j = nbr(iz, -m, L.nz, L.pz);
    if (j >= 0) s = zadd(s, zsc(L.cz[m], zmul(zcj(L.ez[m]), at(ix, iy, j))));
  }
  for (int m = 1; m <= L.nf; m++) {
    const int j = nbr(iz, m, L.nz, L.pz);
    if (j >= 0) s = zadd(s, zsc(L.cz[m], zmul(L.ez[m], at(ix, iy, j))));
  }
  out[ldo * b + p] = s;
}

// K u = 1/2 u + 1/12 sum of the six neighbours (ksprecondition8f.F90:42-75 order), wrapped or
// zero-padded per axis
__global__ void k_zprecond(int nx, int ny, int nz, int px, int py, int pz, const zt *u, long ldu,
                           zt *out, long ldo) {
  const long n = (long)nx * ny * nz;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  u += ldu * blockIdx.y;
  out += ldo * blockIdx.y;
  const int ix = (int)(p % nx), iy = (int)((p / nx) % ny), iz = (int)(p / ((long)nx * ny));
  const double w = 1.0 / 12.0;
  auto at = [&](int x, int y, int z) {
    return (x < 0 || y < 0 || z < 0) ? zmk(0.0, 0.0) : u[x + (long)nx * (y + (long)ny * z)];
  };
  zt s = zsc(w, at(nbr(ix, -1, nx, px), iy, iz));
  s = zadd(s, zsc(w, at(ix, nbr(iy, -1, ny, py), iz)));
  s = zadd(s, zsc(w, at(ix, iy, nbr(iz, -1, nz, pz))));
  s = zadd(s, zsc(0.5, u[p]));
  s = zadd(s, zsc(w, at(ix, iy, nbr(iz, 1, nz, pz))));
  s = zadd(s, zsc(w, at(ix, nbr(iy, 1, ny, py), iz)));
  s = zadd(s, zsc(w, at(nbr(ix, 1, nx, px), iy, iz)));
  out[p] = s;
}

// ----------------------------------------------------------------------------- band arithmetic
// res_b = scale sum_i conj(x_b(i)) y_b(i), deterministic two-stage
constexpr int ZD_T = 256;
__global__ __launch_bounds__(ZD_T) void k_zdot_part(long n, const zt *x, long ldx, const zt *y,
                                                    long ldy, zt *part, int np) {
  __shared__ zt red[ZD_T];
  const int b = blockIdx.y, t = threadIdx.x;
  const long chunk = (n + np - 1) / np, i0 = chunk * blockIdx.x, i1 = min(n, i0 + chunk);
  zt s = zmk(0.0, 0.0);
  for (long i = i0 + t; i < i1; i += ZD_T) s = zadd(s, zmul(zcj(x[ldx * b + i]), y[ldy * b + i]));
  red[t] = s;
  __syncthreads();
  for (int w = ZD_T / 2; w > 0; w >>= 1) {
    if (t < w) red[t] = zadd(red[t], red[t + w]);
    __syncthreads();
  }
  if (t == 0) part[(long)np * b + blockIdx.x] = red[0];
}
__global__ void k_zdot_red(const zt *part, int np, double scale, zt *res) {
  const int b = blockIdx.x;
  if (threadIdx.x) return;
  zt s = zmk(0.0, 0.0);
  for (int i = 0; i < np; i++) s = zadd(s, part[(long)np * b + i]);
  res[b] = zsc(scale, s);
}

// out = a x + c y (complex scalars; y may be null)
__global__ void k_zaxpby(long n, zt a, const zt *x, zt c, const zt *y, zt *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  zt v = zmul(a, x[i]);
  if (y) v = zadd(v, zmul(c, y[i]));
  out[i] = v;
}
// per-band real scale
__global__ void k_zscale(long n, double s, zt *x) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = zsc(s, x[i]);
}
// per-band coefficients: out_b = ca[b] x_b + cb[b] y_b (y may be null); out may alias x or y
__global__ void k_zcomb(long n, const zt *ca, const zt *x, long ldx, const zt *cb, const zt *y,
                        long ldy, zt *out, long ldo) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= n) return;
  zt v = zmul(ca[b], x[ldx * b + i]);
  if (y) v = zadd(v, zmul(cb[b], y[ldy * b + i]));
  out[ldo * b + i] = v;
}

// rho += sum_b f_b |u_b|^2 (band order as scf_charge8f.F90:354-362)
__global__ void k_zcharge(long n, int nb, const zt *u, long ldu, const double *f, double *rho) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = rho[i];
  for (int b = 0; b < nb; b++) {
    const zt v = u[ldu * b + i];
    r = r + f[b] * (v.x * v.x + v.y * v.y);
  }
  rho[i] = r;
}

// ----------------------------------------------------------------------------- halo'd leaves
struct ZHalo {
  int na, nb, nc, nf, nf1;
  long ex, ey, ez;
};
__device__ inline long hx(const ZHalo &h, int ia, int ib, int ic) {
  return (long)(ia + h.nf1) + h.ex * ((long)(ib + h.nf1) + h.ey * (long)(ic + h.nf1));
}
// periodic wrap of nsnd face layers along one axis (overlap_finitedifference8f.F90:327-645
// and :830-1010 on one rank); component is = blockIdx.y
__global__ void k_zhalo_axis(ZHalo h, int axis, int nsnd, zt *v) {
  zt *w = v + (long)blockIdx.y * h.ex * h.ey * h.ez;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (axis == 0) {
    const long cnt = (long)nsnd * h.nb * h.nc;
    if (q >= cnt) return;
    const int iov = (int)(q % nsnd) + 1, ib = (int)((q / nsnd) % h.nb) + 1, ic = (int)(q / ((long)nsnd * h.nb)) + 1;
    const zt lo = w[hx(h, h.na - nsnd + iov, ib, ic)], hi = w[hx(h, iov, ib, ic)];
    w[hx(h, -nsnd + iov, ib, ic)] = lo;
    w[hx(h, h.na + iov, ib, ic)] = hi;
  } else if (axis == 1) {
    const long cnt = (long)h.na * nsnd * h.nc;
    if (q >= cnt) return;
    const int ia = (int)(q % h.na) + 1, iov = (int)((q / h.na) % nsnd) + 1, ic = (int)(q / ((long)h.na * nsnd)) + 1;
    const zt lo = w[hx(h, ia, h.nb - nsnd + iov, ic)], hi = w[hx(h, ia, iov, ic)];
    w[hx(h, ia, -nsnd + iov, ic)] = lo;
    w[hx(h, ia, h.nb + iov, ic)] = hi;
  } else {
    const long cnt = (long)h.na * h.nb * nsnd;
    if (q >= cnt) return;
    const int ia = (int)(q % h.na) + 1, ib = (int)((q / h.na) % h.nb) + 1, iov = (int)(q / ((long)h.na * h.nb)) + 1;
    const zt lo = w[hx(h, ia, ib, h.nc - nsnd + iov)], hi = w[hx(h, ia, ib, iov)];
    w[hx(h, ia, ib, -nsnd + iov)] = lo;
    w[hx(h, ia, ib, h.nc + iov)] = hi;
  }
}

// vcm (halo'd) *= e^{i k.r} with r from the unwrapped index (kslaplacian8f.F90:578-592)
__global__ void k_zphase_halo(ZHalo h, double dx, double dy, double dz, double xmax, double ymax,
                              double zmax, double kx, double ky, double kz, zt *v) {
  const long tot = h.ex * h.ey * h.ez;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= tot) return;
  const int ix = (int)(q % h.ex) - h.nf1, iy = (int)((q / h.ex) % h.ey) - h.nf1,
            iz = (int)(q / (h.ex * h.ey)) - h.nf1;
  const double x = ix * dx - xmax - 0.5 * dx, y = iy * dy - ymax - 0.5 * dy, z = iz * dz - zmax - 0.5 * dz;
  const double th = kx * x + ky * y + kz * z;
  double s, c;
  sincos(th, &s, &c);
  v[q] = zmul(zmk(c, s), v[q]);
}

// workc = (-1/2 lap_nf + veff) of the halo'd (phased) input, kslaplacian_c_01's three sweeps
__global__ void k_zlap_halo(ZHalo h, double c0, ZLap L, const zt *v, const double *veff, zt *w) {
  const long n = (long)h.na * h.nb * h.nc;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % h.na) + 1, iy = (int)((p / h.na) % h.nb) + 1, iz = (int)(p / ((long)h.na * h.nb)) + 1;
  zt s = zmk(0.0, 0.0);
  for (int m = L.nf; m >= 1; m--) s = zadd(s, zsc(L.cx[m], v[hx(h, ix - m, iy, iz)]));
  s = zadd(s, zsc(c0 + veff[p], v[hx(h, ix, iy, iz)]));
  for (int m = 1; m <= L.nf; m++) s = zadd(s, zsc(L.cx[m], v[hx(h, ix + m, iy, iz)]));
  for (int m = L.nf; m >= 1; m--) s = zadd(s, zsc(L.cy[m], v[hx(h, ix, iy - m, iz)]));
  for (int m = 1; m <= L.nf; m++) s = zadd(s, zsc(L.cy[m], v[hx(h, ix, iy + m, iz)]));
  for (int m = L.nf; m >= 1; m--) s = zadd(s, zsc(L.cz[m], v[hx(h, ix, iy, iz - m)]));
  for (int m = 1; m <= L.nf; m++) s = zadd(s, zsc(L.cz[m], v[hx(h, ix, iy, iz + m)]));
  w[p] = s;
}

// avcm += e^{-i k.r} workc on the interior
__global__ void k_zunphase_add(int na, int nb, int nc, double dx, double dy, double dz,
                               double xmax, double ymax, double zmax, double kx, double ky,
                               double kz, const zt *w, zt *out) {
  const long n = (long)na * nb * nc;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % na) + 1, iy = (int)((p / na) % nb) + 1, iz = (int)(p / ((long)na * nb)) + 1;
  const double x = ix * dx - xmax - 0.5 * dx, y = iy * dy - ymax - 0.5 * dy, z = iz * dz - zmax - 0.5 * dz;
  const double th = kx * x + ky * y + kz * z;
  double s, c;
  sincos(th, &s, &c);
  out[p] = zadd(out[p], zmul(zmk(c, -s), w[p]));
}

// K u on a halo'd input (the leaf entry point)
__global__ void k_zprecond_halo(ZHalo h, const zt *v, zt *out) {
  const long n = (long)h.na * h.nb * h.nc;
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % h.na) + 1, iy = (int)((p / h.na) % h.nb) + 1, iz = (int)(p / ((long)h.na * h.nb)) + 1;
  const double w = 1.0 / 12.0;
  zt s = zsc(w, v[hx(h, ix - 1, iy, iz)]);
  s = zadd(s, zsc(w, v[hx(h, ix, iy - 1, iz)]));
  s = zadd(s, zsc(w, v[hx(h, ix, iy, iz - 1)]));
  s = zadd(s, zsc(0.5, v[hx(h, ix, iy, iz)]));
  s = zadd(s, zsc(w, v[hx(h, ix, iy, iz + 1)]));
  s = zadd(s, zsc(w, v[hx(h, ix, iy + 1, iz)]));
  s = zadd(s, zsc(w, v[hx(h, ix + 1, iy, iz)]));
  out[p] = s;
}

ZLap make_lap(const Grid &g, int nf, double kx, double ky, double kz) {
  ZLap L{};
  L.nx = g.nx; L.ny = g.ny; L.nz = g.nz; L.nf = nf;
  L.px = g.nperi >= 1; L.py = g.nperi >= 2; L.pz = g.nperi >= 3;
  double a[9];
  fd_coefficients(nf, a);
  for (int m = 0; m < 9; m++) {
    L.cx[m] = a[m] / g.dx / g.dx * (-0.5);
    L.cy[m] = a[m] / g.dy / g.dy * (-0.5);
    L.cz[m] = a[m] / g.dz / g.dz * (-0.5);
    L.ex[m] = zmk(std::cos(kx * m * g.dx), std::sin(kx * m * g.dx));
    L.ey[m] = zmk(std::cos(ky * m * g.dy), std::sin(ky * m * g.dy));
    L.ez[m] = zmk(std::cos(kz * m * g.dz), std::sin(kz * m * g.dz));
  }
  L.c0 = L.cx[0] + L.cy[0] + L.cz[0];
  return L;
}

}  // namespace

// ============================================================================= ZCtx
double2 *ZCtx::zbuf(const char *tag, size_t count) {
  return (double2 *)c.buf(tag, 2 * std::max<size_t>(1, count));
}

void ZCtx::phases(const int *lstx, const int *lsty, const int *lstz, const int *natx,
                  const int *naty, const int *natz) {
  const NonlocalPlan &pl = c.pl;
  ph = zbuf("z.ph", (size_t)pl.num_list * std::max(1, pl.num_ppcell));
  if (pl.nact == 0) return;
  hipLaunchKernelGGL(k_zphase, dim3(nb_(std::max(1, pl.max_natinf), 256), pl.nact), dim3(256), 0, c.s,
                     pl.act, pl.nact, pl.num_list, lstx, lsty, lstz, natx, naty, natz, c.g.dx, c.g.dy,
                     c.g.dz, kx, ky, kz, ph);
  HIPCK(hipGetLastError());
}

void ZCtx::proj(const zt *u, long ldu, int nb, zt *cp, long ldcp, bool accumulate) {
  const NonlocalPlan &pl = c.pl;
  if (n1() <= 0 || nb <= 0) return;
  if (!accumulate)
    HIPCK(hipMemsetAsync(cp, 0, sizeof(zt) * ((size_t)ldcp * (nb - 1) + (size_t)n1()), c.s));
  if (pl.nact == 0) return;
  if (pl.max_iprj > ZG_J) throw Error(1, "complex path: more than 16 projectors per atom");
  hipLaunchKernelGGL(k_zgather, dim3(pl.nact, nb), dim3(ZG_T), 0, c.s, pl.act, pl.num_list, pl.nprjmx,
                     pl.vnl, pl.lst, ph, u, ldu, cp, ldcp, 1);
  HIPCK(hipGetLastError());
}

void ZCtx::expand(int nhs, int ns, const zt *cp, long ldcp, int nb, zt *out, long ldo) {
  const NonlocalPlan &pl = c.pl;
  if (pl.nact == 0 || nb <= 0 || pl.nnz == 0) return;
  zt *coef = zbuf("z.coef", (size_t)n1() * nb);
  const int d3 = nhs == 1 ? pl.nums : 1;
  hipLaunchKernelGGL(k_zcoef, dim3(pl.nact, nb), dim3(std::max(1, pl.max_iprj)), 0, c.s, pl.act,
                     pl.nact, pl.nprjmx, nhs, d3, nhs == 1 ? ns : 1, nhs == 1 ? c.dij : c.sss, cp, ldcp,
                     1.0 / c.g.dv(), coef, n1());
  hipLaunchKernelGGL(k_zexpand, dim3(nb_(pl.nnz, 256), nb), dim3(256), 0, c.s, pl.nzrows, pl.nnz,
                     pl.row_ptr, pl.ent, pl.num_list, pl.nprjmx, pl.iprj_slot, pl.vnl, ph, coef, n1(),
                     out, ldo);
  HIPCK(hipGetLastError());
}

void ZCtx::hpsi(const double *veff, const zt *u, long ldu, int nb, const zt *cp, long ldcp, zt *out,
                long ldo) {
  if (nb <= 0) return;
  ZLap L = make_lap(c.g, c.g.nf, kx, ky, kz);
  hipLaunchKernelGGL(k_zlap, dim3(nb_(n(), 256), nb), dim3(256), 0, c.s, L, u, ldu, veff, out, ldo, 0);
  HIPCK(hipGetLastError());
  expand(1, ns_h, cp, ldcp, nb, out, ldo);
}

void ZCtx::spsi(const zt *u, long ldu, int nb, const zt *cp, long ldcp, zt *out, long ldo) {
  if (out != u)
    HIPCK(hipMemcpy2DAsync(out, sizeof(zt) * ldo, u, sizeof(zt) * ldu, sizeof(zt) * n(), nb,
                           hipMemcpyDeviceToDevice, c.s));
  if (!c.q_zero) expand(0, 1, cp, ldcp, nb, out, ldo);
}

void ZCtx::precond(const zt *u, zt *out, int nb, long ldu, long ldo) {
  const Grid &g = c.g;
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_zprecond, dim3(nb_(n(), 256), nb), dim3(256), 0, c.s, g.nx, g.ny, g.nz,
                     (int)(g.nperi >= 1), (int)(g.nperi >= 2), (int)(g.nperi >= 3), u,
                     ldu > 0 ? ldu : n(), out, ldo > 0 ? ldo : n());
  HIPCK(hipGetLastError());
}

void ZCtx::dots(const zt *x, long ldx, const zt *y, long ldy, int nb, zt *res, double scale) {
  const long nn = n();
  const int np = (int)std::min<long>(256, std::max<long>(1, (nn + 8191) / 8192));
  zt *part = zbuf("z.dotpart", (size_t)np * nb);
  hipLaunchKernelGGL(k_zdot_part, dim3(np, nb), dim3(ZD_T), 0, c.s, nn, x, ldx, y, ldy, part, np);
  hipLaunchKernelGGL(k_zdot_red, dim3(nb), dim3(64), 0, c.s, part, np, scale, res);
  HIPCK(hipGetLastError());
}

zc ZCtx::dot1(const zt *x, const zt *y, double scale) {
  zt *r = zbuf("z.dot1", 1);
  dots(x, n(), y, n(), 1, r, scale);
  zt h;
  HIPCK(hipMemcpyAsync(&h, r, sizeof(zt), hipMemcpyDeviceToHost, c.s));
  sync();
  return zc(h.x, h.y);
}

void ZCtx::axpby(long len, zc a, const zt *x, zc b, const zt *y, zt *out) {
  hipLaunchKernelGGL(k_zaxpby, dim3(nb_(len, 256)), dim3(256), 0, c.s, len, zmk(a.real(), a.imag()), x,
                     zmk(b.real(), b.imag()), y, out);
  HIPCK(hipGetLastError());
}

void ZCtx::scale(long len, double s, zt *x) {
  hipLaunchKernelGGL(k_zscale, dim3(nb_(len, 256)), dim3(256), 0, c.s, len, s, x);
  HIPCK(hipGetLastError());
}

void ZCtx::comb(long len, int nb, const zc *ca, const zt *x, long ldx, const zc *cb, const zt *y,
                long ldy, zt *out, long ldo) {
  if (nb <= 0 || len <= 0) return;
  std::vector<zt> h(2 * (size_t)nb);
  for (int b = 0; b < nb; b++) {
    h[b] = zmk(ca[b].real(), ca[b].imag());
    h[nb + b] = cb ? zmk(cb[b].real(), cb[b].imag()) : zmk(0.0, 0.0);
  }
  zt *d = zbuf("z.comb", 2 * (size_t)nb);
  HIPCK(hipMemcpyAsync(d, h.data(), sizeof(zt) * 2 * nb, hipMemcpyHostToDevice, c.s));
  hipLaunchKernelGGL(k_zcomb, dim3(nb_(len, 256), nb), dim3(256), 0, c.s, len, d, x, ldx, d + nb, y,
                     ldy, out, ldo);
  HIPCK(hipGetLastError());
  sync();  // the host coefficients go out of scope
}

namespace {

// Cholesky-QR of bands [0, M) in the S metric: the Gram-Schmidt Q of
// orthogonalization_c_01 with l_inp = 0 (orthogonalization8f.F90:379-724)
void zcholqr(ZCtx &z, int M, zt *psi, zt *spsi, zt *cps) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  rocblas_handle h = lib().blas;
  zt *G = z.zbuf("z.G", (size_t)M * M);
  const rocblas_double_complex dv(c.g.dv(), 0.0), zero(0.0, 0.0), one(1.0, 0.0);
  BLASCK(rocblas_zgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, M, M, (int)n,
                       &dv, (const rocblas_double_complex *)psi, (int)n,
                       (const rocblas_double_complex *)spsi, (int)n, &zero, (rocblas_double_complex *)G, M));
  int *info = (int *)scratch("z.info", sizeof(int));
  BLASCK(rocsolver_zpotrf(h, rocblas_fill_upper, M, (rocblas_double_complex *)G, M, info));
  int hinfo = 0;
  HIPCK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.s));
  sync();
  if (hinfo != 0) throw Error(2, "orthogonalization_c_01 (1): wf has negative norm");
  auto tr = [&](zt *X, long ld, long rows) {
    BLASCK(rocblas_ztrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                         rocblas_diagonal_non_unit, (int)rows, M, &one, (const rocblas_double_complex *)G,
                         M, (rocblas_double_complex *)X, (int)ld));
  };
  tr(psi, n, n);
  if (spsi != psi) tr(spsi, n, n);
  if (n1 > 0) tr(cps, n1, n1);
}

// band l (0-based) against bands 0..l-1 in the S metric, then S-normalised
// (orthogonalization8f.F90:628-716)
void zgs_one(ZCtx &z, int l, zt *psi, zt *spsi, zt *cps, const char *msg) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  rocblas_handle h = lib().blas;
  zt *v = psi + n * l, *sv = spsi + n * l, *cv = cps + n1 * l;
  if (l > 0) {
    zt *av = z.zbuf("z.gs1", l);
    const rocblas_double_complex dv(c.g.dv(), 0.0), zero(0.0, 0.0), one(1.0, 0.0), mone(-1.0, 0.0);
    BLASCK(rocblas_zgemv(h, rocblas_operation_conjugate_transpose, (int)n, l, &dv,
                         (const rocblas_double_complex *)spsi, (int)n, (const rocblas_double_complex *)v,
                         1, &zero, (rocblas_double_complex *)av, 1));
    BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n, l, &mone, (const rocblas_double_complex *)psi,
                         (int)n, (const rocblas_double_complex *)av, 1, &one, (rocblas_double_complex *)v, 1));
    if (spsi != psi)
      BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n, l, &mone,
                           (const rocblas_double_complex *)spsi, (int)n,
                           (const rocblas_double_complex *)av, 1, &one, (rocblas_double_complex *)sv, 1));
    if (n1 > 0)
      BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n1, l, &mone,
                           (const rocblas_double_complex *)cps, (int)n1,
                           (const rocblas_double_complex *)av, 1, &one, (rocblas_double_complex *)cv, 1));
  }
  const double vecnor = z.dot1(v, sv, c.g.dv()).real();
  if (vecnor <= 0.0) throw Error(2, msg);
  const double t = 1.0 / std::sqrt(vecnor);
  z.scale(n, t, v);
  if (sv != v) z.scale(n, t, sv);
  if (n1 > 0) z.scale(n1, t, cv);
}

}  // namespace

// orthogonalization_c_01 for one spin / k-point: bands [l0, l1) (all when l_inp = 0)
void zortho_c_01(ZCtx &z, int M, int l_inp, int nflag, int key_cmpt, zt *psi, zt *spsi, zt *cps) {
  const long n = z.n(), n1 = z.n1();
  const int l0 = l_inp == 0 ? 0 : l_inp - 1, l1 = l_inp == 0 ? M : l_inp;
  if (nflag == key_cmpt) {
    z.proj(psi + n * l0, n, l1 - l0, cps + n1 * l0, n1);
    z.spsi(psi + n * l0, n, l1 - l0, cps + n1 * l0, n1, spsi + n * l0, n);
  }
  if (l_inp == 0) zcholqr(z, M, psi, spsi, cps);
  else zgs_one(z, l_inp - 1, psi, spsi, cps, "orthogonalization_c_01 (3): wf has negative norm");
}

// rayleighritz_c (rayleigh-ritz8f.F90:128-242): H Psi (l_rrb), A = dV Psi^H H Psi, zheevd
// (upper), Psi <- Psi Z, S Psi <- S Psi Z (hsvcm ends as S Psi Z), cspsep <- cspsep Z
void zrayleigh_ritz(ZCtx &z, int M, const double *veff, bool l_rrb, zt *psi, zt *spsi, zt *cps,
                    zt *hpsi, double *sval) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  rocblas_handle h = lib().blas;
  if (l_rrb) z.hpsi(veff, psi, n, M, cps, n1, hpsi, n);
  zt *A = z.zbuf("z.rrA", (size_t)M * M);
  const rocblas_double_complex dv(c.g.dv(), 0.0), zero(0.0, 0.0), one(1.0, 0.0);
  BLASCK(rocblas_zgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, M, M, (int)n,
                       &dv, (const rocblas_double_complex *)psi, (int)n,
                       (const rocblas_double_complex *)hpsi, (int)n, &zero, (rocblas_double_complex *)A, M));
  double *E = c.buf("z.rrE", M), *work = c.buf("z.rrwork", M);
  int *info = (int *)scratch("z.info", sizeof(int));
  BLASCK(rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, M, (rocblas_double_complex *)A, M,
                          E, work, info));
  int hinfo = 0;
  HIPCK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c.s));
  sync();
  if (hinfo != 0) throw Error(3, "rayleighritz_diag_c: error in matrix diagonalization");
  HIPCK(hipMemcpyAsync(sval, E, sizeof(double) * M, hipMemcpyDeviceToDevice, c.s));
  auto rot = [&](zt *X, long ld, long rows, zt *out) {
    BLASCK(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (int)rows, M, M, &one,
                         (const rocblas_double_complex *)X, (int)ld, (const rocblas_double_complex *)A, M,
                         &zero, (rocblas_double_complex *)out, (int)ld));
  };
  rot(psi, n, n, hpsi);
  HIPCK(hipMemcpyAsync(psi, hpsi, sizeof(zt) * n * M, hipMemcpyDeviceToDevice, c.s));
  if (spsi != psi) {
    rot(spsi, n, n, hpsi);
    HIPCK(hipMemcpyAsync(spsi, hpsi, sizeof(zt) * n * M, hipMemcpyDeviceToDevice, c.s));
  }
  if (n1 > 0) {
    zt *tc = z.zbuf("z.rrcp", (size_t)n1 * M);
    rot(cps, n1, n1, tc);
    HIPCK(hipMemcpyAsync(cps, tc, sizeof(zt) * n1 * M, hipMemcpyDeviceToDevice, c.s));
  }
}

// kscg_c band by band (kscg8f.F90:492-1032; per-band scalars on the host)
void zkscg_c(ZCtx &z, int M, int northo, int nkscg, int nsdmax, int nprecon_cg, int key_cmpt,
             int key_nocmpt, double epssd, const double *veff, zt *Psi, zt *SPsi, zt *Cps, zt *H,
             double *sval, double *resid, int *ksconv, int *ksitmax) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  const double dv = c.g.dv();
  rocblas_handle h = lib().blas;
  zt *rvec = z.zbuf("zcg.rvec", n), *pvec = z.zbuf("zcg.pvec", n), *pv = z.zbuf("zcg.pv", n);
  zt *rv = z.zbuf("zcg.rv", n), *spv = z.zbuf("zcg.spv", n), *av = z.zbuf("zcg.av", n);
  zt *asv = z.zbuf("zcg.asv", n), *avc = z.zbuf("zcg.avc", M);
  zt *crp = z.zbuf("zcg.crp", n1), *cpp = z.zbuf("zcg.cpp", n1), *cp = z.zbuf("zcg.cp", n1);
  const bool same_s = c.q_zero && SPsi == Psi;
  const int key = northo == 2 ? key_cmpt : key_nocmpt;
  const rocblas_double_complex rdv(dv, 0.0), zero(0.0, 0.0), one(1.0, 0.0), mone(-1.0, 0.0);
  // eigenvalue and residual estimates are host scalars here: collected on the host, stored once
  std::vector<double> hsv(M), hres(M);
  HIPCK(hipMemcpyAsync(hsv.data(), sval, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  HIPCK(hipMemcpyAsync(hres.data(), resid, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  sync();
  auto zero_n = [&](zt *x, long len) { HIPCK(hipMemsetAsync(x, 0, sizeof(zt) * std::max<long>(1, len), c.s)); };
  auto copy = [&](zt *dst, const zt *src, long len) {
    if (len > 0) HIPCK(hipMemcpyAsync(dst, src, sizeof(zt) * len, hipMemcpyDeviceToDevice, c.s));
  };
  // x -= Psi(:, 0..l) dV <S psi_j|x> (kscg8f.F90:751-778)
  auto project = [&](int l, zt *x, zt *sx, zt *cx) {
    BLASCK(rocblas_zgemv(h, rocblas_operation_conjugate_transpose, (int)n, l + 1, &rdv,
                         (const rocblas_double_complex *)SPsi, (int)n, (const rocblas_double_complex *)x, 1,
                         &zero, (rocblas_double_complex *)avc, 1));
    BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n, l + 1, &mone, (const rocblas_double_complex *)Psi,
                         (int)n, (const rocblas_double_complex *)avc, 1, &one, (rocblas_double_complex *)x, 1));
    if (sx)
      BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n, l + 1, &mone,
                           (const rocblas_double_complex *)SPsi, (int)n, (const rocblas_double_complex *)avc,
                           1, &one, (rocblas_double_complex *)sx, 1));
    if (cx && n1 > 0)
      BLASCK(rocblas_zgemv(h, rocblas_operation_none, (int)n1, l + 1, &mone,
                           (const rocblas_double_complex *)Cps, (int)n1, (const rocblas_double_complex *)avc,
                           1, &one, (rocblas_double_complex *)cx, 1));
  };
  for (int l = 0; l < M; l++) {
    zt *psi = Psi + n * l, *spsi = SPsi + n * l, *hp = H + n * l, *cps = Cps + n1 * l;
    for (int loop = 0; loop < northo; loop++) zortho_c_01(z, M, l + 1, key, key_cmpt, Psi, SPsi, Cps);
    copy(cp, cps, n1);
    z.hpsi(veff, psi, n, 1, cp, n1, hp, n);
    int iiter = 0;
    bool conv = false;
    zc r1r, r2r(0.0, 0.0);
    while (iiter < nsdmax && !conv) {
      iiter++;
      bool doreset = (iiter == 1) || (iiter == 10) || ((iiter - 1) % 200 == 0);
      if (doreset) {
        r2r = zc(1.0, 0.0);
        zero_n(pvec, n); zero_n(rvec, n); zero_n(pv, n); zero_n(spv, n);
        if (n1 > 0) zero_n(cpp, n1);
      }
      r1r = r2r;
      const double valre = z.dot1(psi, hp, dv).real();
      hsv[l] = valre;
      z.axpby(n, zc(-1.0, 0.0), hp, zc(valre, 0.0), spsi, rvec);  // r = -(H psi - eps S psi)
      if (nprecon_cg == 1) z.precond(rvec, rv);
      else copy(rv, rvec, n);
      for (int loop = 0; loop < northo; loop++) project(l, rv, nullptr, nullptr);
      z.proj(rv, n, 1, cp, n1);
      copy(crp, cp, n1);
      if (same_s) copy(asv, rv, n);
      else z.spsi(rv, n, 1, crp, n1, asv, n);
      r2r = z.dot1(rv, rvec, dv);  // sum conj(K r) r (orthogonalization_c_04(v1 = r, v2 = K r))
      const double res = std::abs(r2r);
      hres[l] = res;
      if (nkscg == 1) {
        const double a2 = std::abs(r1r);
        const zc t = std::conj(r1r) * r2r / (a2 * a2);  // kscg_c_03
        z.axpby(n, zc(1.0, 0.0), rv, t, pv, pv);
        copy(pvec, pv, n);
        z.axpby(n, zc(1.0, 0.0), asv, t, spv, spv);
        if (n1 > 0) z.axpby(n1, zc(1.0, 0.0), crp, r2r / r1r, cpp, cpp);
      } else {
        copy(pvec, rv, n);
        copy(spv, asv, n);
        copy(cpp, crp, n1);
      }
      copy(cp, cpp, n1);
      copy(asv, spv, n);
      for (int loop = 0; loop < northo; loop++) project(l, pvec, asv, cp);
      const double vecnor = z.dot1(pvec, asv, dv).real();
      if (vecnor <= 0.0) throw Error(2, "kscg_c: pvec has negative norm");
      const double tn = 1.0 / std::sqrt(vecnor);
      z.scale(n, tn, pvec);
      z.scale(n, tn, asv);
      if (n1 > 0) z.scale(n1, tn, cp);
      z.hpsi(veff, pvec, n, 1, cp, n1, av, n);
      const double pap = z.dot1(pvec, av, dv).real();
      const zc sap = z.dot1(psi, av, dv);
      // kscg_c_05 (kscg8f.F90:1233-1274)
      const double eps = 1.0e-14, s2 = std::norm(sap);
      const double gamma = (valre + pap) * 0.5 - std::sqrt((valre - pap) * (valre - pap) * 0.25 + s2);
      double alph;
      zc cbeta;
      if (std::fabs((valre - gamma) / gamma) > eps) {
        const double d = std::sqrt(s2 + (valre - gamma) * (valre - gamma));
        const zc calph = sap / d;
        cbeta = zc(-(valre - gamma) / d, 0.0);
        alph = std::abs(calph);
        if (alph > eps) cbeta = cbeta * alph / calph;
      } else {
        alph = 1.0;
        cbeta = zc(0.0, 0.0);
      }
      z.axpby(n, zc(alph, 0.0), psi, cbeta, pvec, psi);
      z.axpby(n, zc(alph, 0.0), hp, cbeta, av, hp);
      if (!same_s) z.axpby(n, zc(alph, 0.0), spsi, cbeta, asv, spsi);
      if (n1 > 0) z.axpby(n1, zc(alph, 0.0), cps, cbeta, cp, cps);
      conv = std::sqrt(res) < epssd;
      *ksconv = *ksconv && (conv || iiter < nsdmax);
    }
    *ksitmax = std::max(*ksitmax, iiter);
  }
  HIPCK(hipMemcpyAsync(sval, hsv.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  HIPCK(hipMemcpyAsync(resid, hres.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  sync();
}

namespace {

// eigenpairs of the Hermitian 2 x 2 [[a, b], [conj(b), d]] (one complex Jacobi rotation:
// the phase of b is moved to the second component, then the real symmetric rotation),
// eigenvalues in descending order, eigenvectors as columns v[0..1] (v[j][i] = component i)
void herm2(double a, zc b, double d, double w[2], zc v[2][2]) {
  const double ab = std::abs(b);
  if (ab == 0.0) {
    const bool sw = d > a;
    w[0] = sw ? d : a;
    w[1] = sw ? a : d;
    v[0][0] = sw ? 0.0 : 1.0; v[0][1] = sw ? 1.0 : 0.0;
    v[1][0] = sw ? 1.0 : 0.0; v[1][1] = sw ? 0.0 : 1.0;
    return;
  }
  const zc e = std::conj(b) / ab;  // D = diag(1, conj(b)/|b|) makes the off-diagonal |b|
  const double tau = (d - a) / (2.0 * ab);
  const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
  const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
  const double l1 = a - t * ab, l2 = d + t * ab;  // eigenvectors (cs, -sn) and (sn, cs)
  zc v1[2] = {cs, -sn * e}, v2[2] = {sn, cs * e};
  if (l1 >= l2) {
    w[0] = l1; w[1] = l2;
    v[0][0] = v1[0]; v[0][1] = v1[1]; v[1][0] = v2[0]; v[1][1] = v2[1];
  } else {
    w[0] = l2; w[1] = l1;
    v[0][0] = v2[0]; v[0][1] = v2[1]; v[1][0] = v1[0]; v[1][1] = v1[1];
  }
}

}  // namespace

// ksdiis_c (ksdiis8f.F90:715-1509) for one spin / k-point.  With nitermax = 2 every sweep takes
// one DIIS step per band from its incoming vector phi_0 and the trial phi_1 = phi_0 + lambda K r:
// the bands of a block go through each stage together (H, K, projections and all band scalars
// batched), the 2 x 2 subspace problems are solved on the host.  The combination's overall phase
// (zheev's choice in the reference) is fixed by making the incoming band's coefficient real and
// non-negative.
void zksdiis_c(ZCtx &z, const ZDiisArgs &a, const double *veff, zt *Psi, zt *SPsi, zt *Cps, zt *H,
               double *sval, double *resid, int *ksconv, int *ksitmax) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  const int M = a.M;
  const double dv = c.g.dv();
  // bands per block: 7 complex grid vectors each, at most about 8 GiB of them
  const long per_band = 7 * n * (long)sizeof(zt);
  const int nbk = (int)std::max<long>(1, std::min<long>(M, (8L << 30) / std::max<long>(1, per_band)));
  zt *PH1 = z.zbuf("zd.ph1", (size_t)n * nbk), *HPH1 = z.zbuf("zd.hph1", (size_t)n * nbk);
  zt *SPH1 = z.zbuf("zd.sph1", (size_t)n * nbk), *R0 = z.zbuf("zd.r0", (size_t)n * nbk);
  zt *RV = z.zbuf("zd.rv", (size_t)n * nbk), *AV = z.zbuf("zd.av", (size_t)n * nbk);
  zt *ASV = z.zbuf("zd.asv", (size_t)n * nbk);
  zt *CP = z.zbuf("zd.cp", (size_t)n1 * nbk), *CPH1 = z.zbuf("zd.cph1", (size_t)n1 * nbk);
  zt *dd = z.zbuf("zd.dots", (size_t)6 * nbk);
  std::vector<double> hsv(M), hres(M);
  HIPCK(hipMemcpyAsync(hsv.data(), sval, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  HIPCK(hipMemcpyAsync(hres.data(), resid, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  sync();
  std::vector<zt> hd(6 * (size_t)nbk);
  auto fetch = [&](int cnt) {
    HIPCK(hipMemcpyAsync(hd.data(), dd, sizeof(zt) * cnt, hipMemcpyDeviceToHost, c.s));
    sync();
  };
  std::vector<zc> one(nbk, zc(1.0, 0.0)), ca(nbk), cb(nbk);
  for (int jiter = 1; jiter < a.ndiismax; jiter++) {  // ksdiis8f.F90:819-821
    for (int l0 = 0; l0 < M; l0 += nbk) {
      const int nb = std::min(nbk, M - l0);
      zt *P = Psi + n * l0, *SP = SPsi + n * l0, *C = Cps + n1 * l0, *Hh = H + n * l0;
      // eigenvalue and residual (ksdiis8f.F90:868-917): H phi_0 with the stored <p|phi_0>
      z.hpsi(veff, P, n, nb, C, n1, Hh, n);
      z.dots(P, n, Hh, n, nb, dd, dv);
      fetch(nb);
      for (int b = 0; b < nb; b++) {
        hsv[l0 + b] = hd[b].x;
        cb[b] = zc(-hd[b].x, 0.0);
      }
      z.comb(n, nb, one.data(), Hh, n, cb.data(), SP, n, R0, n);
      // K r, <p|K r>, H K r, S K r (:920-1063)
      if (a.nprecon_diis == 1) z.precond(R0, RV, nb, n, n);
      else HIPCK(hipMemcpyAsync(RV, R0, sizeof(zt) * n * nb, hipMemcpyDeviceToDevice, c.s));
      z.proj(RV, n, nb, CP, n1);
      z.hpsi(veff, RV, n, nb, CP, n1, AV, n);
      z.spsi(RV, n, nb, CP, n1, ASV, n);
      // step length from the trial step alambda_diis (:1065-1096)
      for (int b = 0; b < nb; b++) ca[b] = zc(a.alambda_diis, 0.0);
      z.comb(n, nb, one.data(), P, n, ca.data(), RV, n, PH1, n);
      z.comb(n, nb, one.data(), Hh, n, ca.data(), AV, n, HPH1, n);
      z.comb(n, nb, one.data(), SP, n, ca.data(), ASV, n, SPH1, n);
      z.dots(RV, n, Hh, n, nb, dd, 1.0);
      z.dots(RV, n, SP, n, nb, dd + nb, 1.0);
      z.dots(PH1, n, HPH1, n, nb, dd + 2 * nb, 1.0);
      z.dots(PH1, n, SPH1, n, nb, dd + 3 * nb, 1.0);
      fetch(4 * nb);
      std::vector<double> lam(nb);
      for (int b = 0; b < nb; b++) {
        const double sv = hsv[l0 + b], a1 = a.alambda_diis;
        const double s1 = hd[b].x, s2 = hd[nb + b].x, s3 = hd[2 * nb + b].x, s4 = hd[3 * nb + b].x;
        const double aaa = 2.0 * (s1 - sv * s2) * dv;
        const double bbb = (s3 / s4 - sv - aaa * a1) / (a1 * a1);
        double al = 0.5 * aaa / bbb;
        if (al > a.alambda_max) al = a.alambda_max;
        if (al < a.alambda_min) al = a.alambda_min;
        lam[b] = al;
        ca[b] = zc(al, 0.0);
      }
      // trial phi (:1099-1116) and the residual of phi_1 (:1140-1149; r_0 is unchanged)
      z.comb(n, nb, one.data(), P, n, ca.data(), RV, n, PH1, n);
      z.comb(n, nb, one.data(), Hh, n, ca.data(), AV, n, HPH1, n);
      z.comb(n, nb, one.data(), SP, n, ca.data(), ASV, n, SPH1, n);
      z.comb(n1, nb, one.data(), C, n1, ca.data(), CP, n1, CPH1, n1);
      for (int b = 0; b < nb; b++) cb[b] = zc(-hsv[l0 + b], 0.0);
      z.comb(n, nb, one.data(), HPH1, n, cb.data(), SPH1, n, RV, n);  // r_1 (K r no longer needed)
      // DIIS matrices (:1154-1199)
      z.dots(P, n, SP, n, nb, dd, dv);
      z.dots(P, n, SPH1, n, nb, dd + nb, dv);
      z.dots(PH1, n, SPH1, n, nb, dd + 2 * nb, dv);
      z.dots(R0, n, R0, n, nb, dd + 3 * nb, dv);
      z.dots(R0, n, RV, n, nb, dd + 4 * nb, dv);
      z.dots(RV, n, RV, n, nb, dd + 5 * nb, dv);
      fetch(6 * nb);
      for (int b = 0; b < nb; b++) {
        auto g = [&](int k) { return zc(hd[k * nb + b].x, hd[k * nb + b].y); };
        const double tmp = 1.0 / std::sqrt(g(2).real());
        // csm, crm with row and column 2 scaled by tmp (the normalisation of phi_1)
        const double s00 = g(0).real(), s11 = g(2).real() * tmp * tmp;
        const zc s01 = g(1) * tmp;
        const double r00 = g(3).real(), r11 = g(5).real() * tmp * tmp;
        const zc r01 = g(4) * tmp;
        double w[2];
        zc v[2][2];
        herm2(s00, s01, s11, w, v);  // descending
        zc cv[2][2] = {{v[0][0], v[0][1]}, {v[1][0], v[1][1]}};
        const double ev[2] = {w[0], w[1]};
        // overlap eigenvalues below eps_eig_diis are dropped (:1227-1240); with two of them in
        // descending order that leaves the count of those at or above it
        const int nmat = (ev[0] >= a.eps_eig_diis) + (ev[1] >= a.eps_eig_diis);
        zc c0, c1;
        if (nmat > 1) {
          for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) cv[j][i] /= std::sqrt(ev[j]);
          // projected residual matrix cr2 = cv^H crm cv (:1242-1260)
          const zc crm[2][2] = {{r00, r01}, {std::conj(r01), r11}};  // crm[i][k] = row i, col k
          zc t1[2][2], cr2[2][2];
          for (int kk = 0; kk < 2; kk++)
            for (int j = 0; j < 2; j++) t1[kk][j] = crm[kk][0] * cv[j][0] + crm[kk][1] * cv[j][1];
          for (int kk = 0; kk < 2; kk++)
            for (int j = 0; j < 2; j++)
              cr2[kk][j] = std::conj(cv[kk][0]) * t1[0][j] + std::conj(cv[kk][1]) * t1[1][j];
          double w2[2];
          zc v2[2][2];
          herm2(cr2[0][0].real(), cr2[0][1], cr2[1][1].real(), w2, v2);
          hres[l0 + b] = w2[1];
          // the lowest: coefficients on phi_0, phi_1 (:1293-1310), S-normalised with csm0
          const zc x0 = cv[0][0] * v2[1][0] + cv[1][0] * v2[1][1];
          const zc x1 = cv[0][1] * v2[1][0] + cv[1][1] * v2[1][1];
          const double t2 = (std::conj(x0) * s00 * x0 + std::conj(x0) * s01 * x1 +
                             std::conj(x1) * std::conj(s01) * x0 + std::conj(x1) * s11 * x1).real();
          c0 = x0 / std::sqrt(t2);
          c1 = x1 / std::sqrt(t2);
        } else {
          c0 = 0.0;
          c1 = 1.0;
        }
        const double m0 = std::abs(c0);
        if (m0 > 0.0) {
          const zc ph = std::conj(c0) / m0;
          c0 = zc(m0, 0.0);
          c1 *= ph;
        }
        ca[b] = c0;
        cb[b] = c1 * tmp;  // phi_1 is combined unnormalised
        *ksitmax = std::max(*ksitmax, 1);
        *ksconv = *ksconv && (std::sqrt(hres[l0 + b]) <= a.epssd);
      }
      // new band (:1313-1352)
      z.comb(n, nb, ca.data(), P, n, cb.data(), PH1, n, P, n);
      z.comb(n, nb, ca.data(), SP, n, cb.data(), SPH1, n, SP, n);
      z.comb(n, nb, ca.data(), Hh, n, cb.data(), HPH1, n, Hh, n);
      z.comb(n1, nb, ca.data(), C, n1, cb.data(), CPH1, n1, C, n1);
    }
    // l_hsv: orthonormalise, then H psi of every band (:1416-1478)
    if (a.l_hsv) {
      for (int i = 0; i < a.northo; i++) zortho_c_01(z, M, 0, a.key_cmpt, a.key_cmpt, Psi, SPsi, Cps);
      z.hpsi(veff, Psi, n, M, Cps, n1, H, n);
    }
  }
  HIPCK(hipMemcpyAsync(sval, hsv.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  HIPCK(hipMemcpyAsync(resid, hres.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  sync();
}

// reorderstates with nrc = 1 (reorderstates8f.F90:70-124): stable rank sort by sval
void zreorder(ZCtx &z, int M, double *sval, double *resid, zt *psi, zt *spsi, zt *cps, zt *hpsi) {
  Ctx &c = z.c;
  const long n = z.n(), n1 = z.n1();
  std::vector<double> sv(M), rs(M);
  HIPCK(hipMemcpyAsync(sv.data(), sval, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  HIPCK(hipMemcpyAsync(rs.data(), resid, sizeof(double) * M, hipMemcpyDeviceToHost, c.s));
  sync();
  std::vector<int> lank(M, 1);
  for (int j = 1; j < M; j++)
    for (int i = 0; i < j; i++) {
      if (sv[i] > sv[j]) lank[i]++;
      else lank[j]++;
    }
  std::vector<double> st(M), rt(M);
  for (int j = 0; j < M; j++) {
    st[lank[j] - 1] = sv[j];
    rt[lank[j] - 1] = rs[j];
  }
  HIPCK(hipMemcpyAsync(sval, st.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  HIPCK(hipMemcpyAsync(resid, rt.data(), sizeof(double) * M, hipMemcpyHostToDevice, c.s));
  auto perm = [&](zt *x, long len, zt *tmp) {
    for (int j = 0; j < M; j++)
      HIPCK(hipMemcpyAsync(tmp + len * (lank[j] - 1), x + len * j, sizeof(zt) * len,
                           hipMemcpyDeviceToDevice, c.s));
    HIPCK(hipMemcpyAsync(x, tmp, sizeof(zt) * len * M, hipMemcpyDeviceToDevice, c.s));
  };
  if (n1 > 0) perm(cps, n1, z.zbuf("z.reo", (size_t)n1 * M));
  perm(psi, n, hpsi);  // hsvcm is the work array, as in the reference
  if (spsi != psi) perm(spsi, n, hpsi);
  sync();
}

void zcharge(long n, int nb, const void *u, long ldu, const double *f, double *rho, hipStream_t s) {
  hipLaunchKernelGGL(k_zcharge, dim3(nb_(n, 256)), dim3(256), 0, s, n, nb, (const zt *)u, ldu, f, rho);
  HIPCK(hipGetLastError());
}

// ----------------------------------------------------------------------------- leaves
void zhalo_fill(int nperi, int na, int nb, int nc, int nf, int nf1, int nsnd, int ncol, void *v,
                hipStream_t s) {
  ZHalo h{na, nb, nc, nf, nf1, (long)na + nf + nf1 + 1, (long)nb + nf + nf1 + 1, (long)nc + nf + nf1 + 1};
  const long cnt[3] = {(long)nsnd * nb * nc, (long)na * nsnd * nc, (long)na * nb * nsnd};
  for (int axis = 0; axis < 3; axis++) {
    if (nperi < axis + 1) continue;
    hipLaunchKernelGGL(k_zhalo_axis, dim3(nb_(cnt[axis], 256), ncol), dim3(256), 0, s, h, axis, nsnd, (zt *)v);
    HIPCK(hipGetLastError());
  }
}

void zlap_leaf(const Grid &g, int nf, double xmax, double ymax, double zmax, double kx, double ky,
               double kz, void *vcm, const double *veff, void *avcm_l, void *workc, hipStream_t s) {
  ZHalo h{g.nx, g.ny, g.nz, nf, nf - 1, (long)g.nx + 2 * nf, (long)g.ny + 2 * nf, (long)g.nz + 2 * nf};
  ZLap L = make_lap(g, nf, 0.0, 0.0, 0.0);
  const long tot = h.ex * h.ey * h.ez, n = g.n();
  hipLaunchKernelGGL(k_zphase_halo, dim3(nb_(tot, 256)), dim3(256), 0, s, h, g.dx, g.dy, g.dz, xmax,
                     ymax, zmax, kx, ky, kz, (zt *)vcm);
  hipLaunchKernelGGL(k_zlap_halo, dim3(nb_(n, 256)), dim3(256), 0, s, h, L.c0, L, (const zt *)vcm, veff,
                     (zt *)workc);
  hipLaunchKernelGGL(k_zunphase_add, dim3(nb_(n, 256)), dim3(256), 0, s, g.nx, g.ny, g.nz, g.dx, g.dy,
                     g.dz, xmax, ymax, zmax, kx, ky, kz, (const zt *)workc, (zt *)avcm_l);
  HIPCK(hipGetLastError());
}

void zprecond_leaf(const Grid &g, int nf, const void *vcm, void *avcm, int ncol, hipStream_t s) {
  ZHalo h{g.nx, g.ny, g.nz, nf, nf - 1, (long)g.nx + 2 * nf, (long)g.ny + 2 * nf, (long)g.nz + 2 * nf};
  const long n = g.n();
  for (int is = 0; is < ncol; is++)
    hipLaunchKernelGGL(k_zprecond_halo, dim3(nb_(n, 256)), dim3(256), 0, s, h,
                       (const zt *)vcm + is * h.ex * h.ey * h.ez, (zt *)avcm + is * n);
  HIPCK(hipGetLastError());
}

}  // namespace kukan
