// force.hip — force() of force8f.F90 on the GPU (declared in include/kukan_hip.h), one rank:
// the local-potential and compensation-charge forces (force_01), the partial-core-correction
// force (force_02), the jellium slab, the core-core interaction, the bias field and the removal
// of the net force.
//
// Every sum runs on the device in a fixed order (per-work-group partial sums, then ordered
// reductions), so the result does not depend on scheduling; it differs from the reference's
// serial sums by rounding only.  Reorderings that leave the mathematics unchanged:
// - force_01's soft part streams dV_loc/dR of all atoms against the coarse density (the
//   dominant traffic: 3 N natom doubles), one pass with the density shared through the cache;
// - the core-correction force visits, for each atom and each of the 27 images the reference
//   sums over, only the dense points of the box around the image's sphere (points outside the
//   sphere contribute nothing), and sums the spins' v_xc at each point before multiplying;
// - the jellium terms take the sum over the slab's Fourier components k_z per z-plane first
//   (it depends on z only), then one pass over the dense grid per atom;
// - the core-core pairs are evaluated one thread per pair, then summed over partners in order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "abi_common.hpp"
#include "kukan_hip.h"

using namespace kukan;
using namespace kukan::abi;

namespace {

constexpr int FT = 256;   // threads per work-group
constexpr int FPT = 16;   // points per thread of the streaming sums

__device__ inline void block_sum3(double &x, double &y, double &z, double *sh) {
  // fixed-order tree over the work-group
  const int t = threadIdx.x;
  sh[t] = x;
  sh[t + FT] = y;
  sh[t + 2 * FT] = z;
  __syncthreads();
  for (int s = FT / 2; s > 0; s >>= 1) {
    if (t < s) {
      sh[t] += sh[t + s];
      sh[t + FT] += sh[t + FT + s];
      sh[t + 2 * FT] += sh[t + 2 * FT + s];
    }
    __syncthreads();
  }
  x = sh[0];
  y = sh[FT];
  z = sh[2 * FT];
}

// force_01, soft part (force8f.F90:2057-2066): partial sums of dV_loc/dR_a(p) rho(p) dV per
// work-group of FT*FPT points; grid.y = atom
__global__ void __launch_bounds__(FT) k_f01_soft(long n, const double *dvx, const double *dvy, const double *dvz,
                                                 const double *rho, double dxyz, double *part, int nblk) {
  __shared__ double sh[3 * FT];
  const int na = blockIdx.y;
  const long base = (long)blockIdx.x * FT * FPT;
  const double *ax = dvx + n * na, *ay = dvy + n * na, *az = dvz + n * na;
  double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll 4
  for (int k = 0; k < FPT; k++) {
    const long p = base + (long)k * FT + threadIdx.x;
    if (p < n) {
      const double r = rho[p];
      x += ax[p] * r * dxyz;
      y += ay[p] * r * dxyz;
      z += az[p] * r * dxyz;
    }
  }
  block_sum3(x, y, z, sh);
  if (threadIdx.x == 0) {
    double *o = part + 3 * ((long)na * nblk + blockIdx.x);
    o[0] = x;
    o[1] = y;
    o[2] = z;
  }
}

struct ListF {
  int num_list_d, key_inps;
  const int *natprid, *napsd, *natinfd, *natinfd_vloc, *lstvecd2, *lstdx, *lstdy, *lstdz, *ndatx, *ndaty, *ndatz;
  const double *hdx, *hdy, *hdz, *drx, *dry, *drz, *rho_aug, *vloc, *vh;
  double ddx, ddy, ddz, ddxyz, biasx, biasy, biasz;
  double *out;  // natom x 3
};

// force_01, hard local part and compensation charge (force8f.F90:2067-2100), one work-group
// per atom
__global__ void __launch_bounds__(FT) k_f01_lists(ListF a) {
  __shared__ double sh[3 * FT];
  const int na = blockIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (a.natprid[na] == a.key_inps) {
    const long col = (long)a.num_list_d * (a.napsd[na] - 1);
    for (int p = threadIdx.x; p < a.natinfd_vloc[na]; p += FT) {
      const double r = a.rho_aug[a.lstvecd2[col + p] - 1];
      x += a.hdx[col + p] * r * a.ddxyz;
      y += a.hdy[col + p] * r * a.ddxyz;
      z += a.hdz[col + p] * r * a.ddxyz;
    }
    const double ox = a.ndatx[na] * a.ddx - 0.5 * a.ddx, oy = a.ndaty[na] * a.ddy - 0.5 * a.ddy,
                 oz = a.ndatz[na] * a.ddz - 0.5 * a.ddz;
    for (int p = threadIdx.x; p < a.natinfd[na]; p += FT) {
      const long i = a.lstvecd2[col + p] - 1;
      const double px = a.lstdx[col + p] * a.ddx + ox, py = a.lstdy[col + p] * a.ddy + oy,
                   pz = a.lstdz[col + p] * a.ddz + oz;
      const double v = a.vloc[i] + a.vh[i] + a.biasx * px + a.biasy * py + a.biasz * pz;
      x += v * a.drx[col + p] * a.ddxyz;
      y += v * a.dry[col + p] * a.ddxyz;
      z += v * a.drz[col + p] * a.ddxyz;
    }
  }
  block_sum3(x, y, z, sh);
  if (threadIdx.x == 0) {
    a.out[3 * na] = x;
    a.out[3 * na + 1] = y;
    a.out[3 * na + 2] = z;
  }
}

// one (atom, image) box of the core-correction force
struct PccBox {
  int na, sp, x0, y0, z0, bx, by, bz;  // 1-based dense start, extents
  double sx, sy, sz;                   // image shift k * 2 xmax
  double thr, dqq, vfac;               // sphere radius, pi / rcutss, 4 pi dqq / (2 pi)^3 / nspv
};

struct PccArgs {
  const PccBox *box;
  const double *coef, *vx, *atx, *aty, *atz;
  const int *nqct;
  int nqmx, nspv, ndx, ndy;
  long nd;
  double xmax, ymax, zmax, ddx, ddy, ddz;
  double *part;  // nbox x nblk x 3
  int nblk;
};

// force_02 (force8f.F90:2141-2183): grid.x = box, grid.y strides the box's points
__global__ void __launch_bounds__(FT) k_pcc(PccArgs a) {
  __shared__ double sh[3 * FT];
  const PccBox &b = a.box[blockIdx.x];
  const long nb = (long)b.bx * b.by * b.bz;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const double *cf = a.coef + (long)(a.nqmx + 1) * (7 + 8L * (b.sp - 1));
  const int nq = a.nqct[b.sp - 1];
  for (long q = (long)blockIdx.y * FT + threadIdx.x; q < nb; q += (long)a.nblk * FT) {
    const int ix = b.x0 + (int)(q % b.bx), iy = b.y0 + (int)((q / b.bx) % b.by), iz = b.z0 + (int)(q / ((long)b.bx * b.by));
    const double x = ix * a.ddx - a.xmax - 0.5 * a.ddx - a.atx[b.na] + b.sx;
    const double y = iy * a.ddy - a.ymax - 0.5 * a.ddy - a.aty[b.na] + b.sy;
    const double z = iz * a.ddz - a.zmax - 0.5 * a.ddz - a.atz[b.na] + b.sz;
    const double r = sqrt(x * x + y * y + z * z);
    if (!(r < b.thr)) continue;
    const double rin = 1.0 / r;
    double t = 0.0;
    for (int iq = 1; iq <= nq; iq++) {
      const double qqq = b.dqq * iq, qqqr = qqq * r, qqqrin = rin / qqq;
      double sn, cn;
      sincos(qqqr, &sn, &cn);
      t += cf[iq] * (cn * rin - sn * qqqrin * rin) * qqq * qqq;
    }
    const long p = (long)(ix - 1) + (long)a.ndx * ((iy - 1) + (long)a.ndy * (iz - 1));
    double v = 0.0;
    for (int ns = 0; ns < a.nspv; ns++) v += a.vx[p + a.nd * ns];
    const double g = t * b.vfac * v * a.ddx * a.ddy * a.ddz * rin;
    fx += g * x;
    fy += g * y;
    fz += g * z;
  }
  block_sum3(fx, fy, fz, sh);
  if (threadIdx.x == 0) {
    double *o = a.part + 3 * ((long)blockIdx.x * a.nblk + blockIdx.y);
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
  }
}

// the jellium slab's sum over k_z != 0 at each dense z-plane (force8f.F90:1746-1770, 1970-1982):
// (4 pi / Omega) chrjel / (strjel - endjel) sum_kz (cos k(strjel - z) - cos k(endjel - z)) / k^2
__global__ void k_jel_planes(int ndz, int nzmax, double ddz, double zmax, double pi, double omega, double chrjel,
                             double strjel, double endjel, double *S) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (iz > ndz) return;
  const double zz = iz * ddz - 0.5 * ddz - zmax;
  double s = 0.0;
  for (int kz = -nzmax; kz <= nzmax; kz++) {
    if (kz == 0) continue;
    const double vkz = pi / zmax * kz, vk2 = vkz * vkz;
    s += 4.0 * pi / omega * chrjel / (strjel - endjel) / vk2 * (cos(vkz * (strjel - zz)) - cos(vkz * (endjel - zz)));
  }
  S[iz - 1] = s;
}

struct JelArgs {
  const double *S, *atx, *aty, *atz, *cp, *rho_aug;
  const int *indspe;
  int ndx, ndy, ndz;
  long nd;
  double xmax, ymax, zmax, ddx, ddy, ddz, pi;
  double *part;  // (natom + 1) x nblk: per atom the slab's force on its core charge; the last
                 // row the slab's force on the augmented density
  int nblk;
};

// jellium part I per atom (grid.y < natom) and the dense part of part II (grid.y == natom)
__global__ void __launch_bounds__(FT) k_jel(JelArgs a, int natom) {
  __shared__ double sh[3 * FT];
  const int na = blockIdx.y;
  double acc = 0.0, d1 = 0.0, d2 = 0.0;
  double cp1 = 0, g2 = 0, g3 = 0, c2 = 0, c3 = 0;
  if (na < natom) {
    const double *c = a.cp + 8L * (a.indspe[na] - 1);
    cp1 = c[0];
    c2 = c[1];
    c3 = c[2];
    g2 = c[3] * pow(c[1] / a.pi, 1.5);
    g3 = c[4] * pow(c[2] / a.pi, 1.5);
  }
  for (long p = (long)blockIdx.x * FT + threadIdx.x; p < a.nd; p += (long)a.nblk * FT) {
    const int ix = (int)(p % a.ndx) + 1, iy = (int)((p / a.ndx) % a.ndy) + 1, iz = (int)(p / ((long)a.ndx * a.ndy)) + 1;
    if (na == natom) {
      acc += a.S[iz - 1] * a.rho_aug[p] * a.ddx * a.ddy * a.ddz;
      continue;
    }
    double x = ix * a.ddx - 0.5 * a.ddx - a.xmax - a.atx[na];
    double y = iy * a.ddy - 0.5 * a.ddy - a.ymax - a.aty[na];
    double z = iz * a.ddz - 0.5 * a.ddz - a.zmax - a.atz[na];
    if (x > a.xmax) x -= 2.0 * a.xmax;
    if (x <= -a.xmax) x += 2.0 * a.xmax;
    if (y > a.ymax) y -= 2.0 * a.ymax;
    if (y <= -a.ymax) y += 2.0 * a.ymax;
    if (z > a.zmax) z -= 2.0 * a.zmax;
    if (z <= -a.zmax) z += 2.0 * a.zmax;
    const double r2 = x * x + y * y + z * z;
    acc += a.S[iz - 1] * cp1 * (g2 * exp(-(c2 * r2)) + g3 * exp(-(c3 * r2))) * a.ddx * a.ddy * a.ddz;
  }
  block_sum3(acc, d1, d2, sh);
  if (threadIdx.x == 0) a.part[(long)na * a.nblk + blockIdx.x] = acc;
}

// jellium part II, one-centre terms (force8f.F90:1944-1965): one work-group per atom in the
// cell, threads over (spin, point, radial index)
struct JelRad {
  const int *natpri, *natpri_inf, *indspe, *nradct;
  const double *radial, *dradial, *point, *wt, *atz, *rhotrur, *rhosmtr, *rhoaugr;
  int key_in, nradmx, npoint, nspv, nzmax;
  double zmax, pi, omega, chrjel, strjel, endjel;
  double *out;  // natom
};
__global__ void __launch_bounds__(FT) k_jel_radial(JelRad a) {
  __shared__ double sh[3 * FT];
  const int na = blockIdx.x;
  double acc = 0.0, d1 = 0.0, d2 = 0.0;
  if (a.natpri[na] == a.key_in) {
    const int ipri = a.natpri_inf[na], sp = a.indspe[na] - 1, nr = a.nradct[sp];
    const int nsp = a.nspv < 2 ? a.nspv : 2;
    const long tot = (long)nsp * a.npoint * (nr - 1);
    for (long u = threadIdx.x; u < tot; u += FT) {
      const int ir = 2 + (int)(u % (nr - 1)), il = (int)((u / (nr - 1)) % a.npoint), ns = (int)(u / ((long)(nr - 1) * a.npoint));
      const long rr = (long)(ir - 1) + (long)a.nradmx * sp;
      const double r = a.radial[rr], dr = a.dradial[rr];
      const double z = a.point[il + 2L * a.npoint] * r + a.atz[na];
      const long ia = (long)(ir - 1) + (long)a.nradmx * (il + (long)a.npoint * (ns + (long)a.nspv * (ipri - 1)));
      const long ib = (long)(ir - 1) + (long)a.nradmx * (il + (long)a.npoint * (ipri - 1));
      const double rho = (a.rhotrur[ia] - a.rhosmtr[ia] - a.rhoaugr[ib]) * r * r * dr * a.wt[il];
      for (int kz = -a.nzmax; kz <= a.nzmax; kz++) {
        if (kz == 0) continue;
        const double vkz = a.pi / a.zmax * kz, vk2 = vkz * vkz;
        acc += 4.0 * a.pi / a.omega * a.chrjel / (a.strjel - a.endjel) / vk2 *
               (cos(vkz * (a.strjel - z)) - cos(vkz * (a.endjel - z))) * rho;
      }
    }
  }
  block_sum3(acc, d1, d2, sh);
  if (threadIdx.x == 0) a.out[na] = acc;
}

struct CoreArgs {
  const double *atx, *aty, *atz, *cp;
  const int *indspe;
  int natom, nperi, nint1dmax, pwx, pwy, pwz, rsx, rsy, rsz;
  double veta, xmax, ymax, zmax, pi, omegain, surf;
  double *pair;  // natom x natom x 3 (l1 fastest)
};

__device__ inline void ewald_real(double cc, double veta, double pi, double x, double y, double z, double &fx,
                                  double &fy, double &fz) {
  const double r = sqrt(x * x + y * y + z * z);
  if (r > 1.0e-20 && (veta * veta) * (r * r) < 100.0) {
    const double vlo0 = -(2.0 * veta / (exp((veta * veta) * (r * r)) * sqrt(pi) * r)) - (1.0 - erf(veta * r)) / (r * r);
    fx -= cc * vlo0 * x / r;
    fy -= cc * vlo0 * y / r;
    fz -= cc * vlo0 * z / r;
  }
}

// the core-core force of atom l1 on atom l0 (force8f.F90:1791-1926), one thread per pair
__global__ void k_core_pairs(CoreArgs a) {
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= (long)a.natom * a.natom) return;
  const int l1 = (int)(u % a.natom), l0 = (int)(u / a.natom);
  const double x0 = a.atx[l0], y0 = a.aty[l0], z0 = a.atz[l0], x1 = a.atx[l1], y1 = a.aty[l1], z1 = a.atz[l1];
  const double cc = a.cp[8L * (a.indspe[l0] - 1)] * a.cp[8L * (a.indspe[l1] - 1)];
  const double pi = a.pi, veta = a.veta;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (a.nperi == 0) {
    const double x = x0 - x1, y = y0 - y1, z = z0 - z1;
    const double r = sqrt(x * x + y * y + z * z);
    if (r > 1.0e-20) {
      const double r3 = r * r * r;
      fx += cc * x / r3;
      fy += cc * y / r3;
      fz += cc * z / r3;
    }
  } else if (a.nperi == 1) {
    const double ta = (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1);
    for (int kx = -a.pwx; kx <= a.pwx; kx++) {
      if (kx == 0) continue;
      const double vkx = pi / a.xmax * kx, tb = vkx * vkx, dt = veta / a.nint1dmax;
      double v0 = 0.0, v1 = 0.0;
      for (int i = 1; i <= a.nint1dmax; i++) {
        const double t = i * dt;
        const double e = exp(-(ta * t * t) - 0.25 * tb / (t * t));
        v0 += e / t * dt;
        v1 += e * t * dt;
      }
      double sn, cn;
      sincos(vkx * (x0 - x1), &sn, &cn);
      const double vcosi = -cn / a.xmax * 0.5, vsine = -sn / a.xmax * 0.5;
      fx -= 2.0 * vkx * cc * vsine * v0;
      fy -= 4.0 * cc * vcosi * v1 * (y0 - y1);
      fz -= 4.0 * cc * vcosi * v1 * (z0 - z1);
    }
    const double vlo0r = ta > 1.0e-16 ? -2.0 * (1.0 - exp(-(veta * veta * ta))) / (ta * 2.0 * a.xmax)
                                      : -(veta * veta / a.xmax);
    fy -= cc * vlo0r * (y0 - y1);
    fz -= cc * vlo0r * (z0 - z1);
    for (int kx = -a.pwx; kx <= a.pwx; kx++)
      ewald_real(cc, veta, pi, kx * 2.0 * a.xmax + x0 - x1, y0 - y1, z0 - z1, fx, fy, fz);
  } else if (a.nperi == 2) {
    for (int ky = -a.pwy; ky <= a.pwy; ky++)
      for (int kx = -a.pwx; kx <= a.pwx; kx++) {
        if (kx == 0 && ky == 0) continue;
        const double vkx = pi / a.xmax * kx, vky = pi / a.ymax * ky;
        const double tb = sqrt(vkx * vkx + vky * vky), ta = z0 - z1;
        const double e1 = 1.0 - erf((tb - 2.0 * veta * veta * ta) / (2.0 * veta));
        const double e2 = 1.0 - erf((tb + 2.0 * veta * veta * ta) / (2.0 * veta));
        const double vep0 = -(pi / tb / a.surf * (e1 / exp(tb * ta) + e2 / exp(-(tb * ta))));
        const double vep1 = -(pi / a.surf * (e1 * exp(-(tb * ta)) - e2 * exp(tb * ta)));
        double sn, cn;
        sincos(vkx * (x0 - x1) + vky * (y0 - y1), &sn, &cn);
        fx -= vkx * cc * (sn * vep0);
        fy -= vky * cc * (sn * vep0);
        fz -= cc * (cn * vep1);
      }
    const double z = fabs(z0 - z1);
    if (z > 1.0e-20) {
      const double vlo1 = -(2.0 * sqrt(pi) / a.surf * cc * (z0 - z1) / z);
      fz -= vlo1 * sqrt(pi) * erf(veta * z);
    }
    for (int ky = -a.rsy; ky <= a.rsy; ky++)
      for (int kx = -a.rsx; kx <= a.rsx; kx++)
        ewald_real(cc, veta, pi, kx * 2.0 * a.xmax + x0 - x1, ky * 2.0 * a.ymax + y0 - y1, z0 - z1, fx, fy, fz);
  } else if (a.nperi == 3) {
    const double eta4 = 4.0 * (veta * veta);
    for (int kz = -a.pwz; kz <= a.pwz; kz++)
      for (int ky = -a.pwy; ky <= a.pwy; ky++)
        for (int kx = -a.pwx; kx <= a.pwx; kx++) {
          if (kx == 0 && ky == 0 && kz == 0) continue;
          const double vkx = pi / a.xmax * kx, vky = pi / a.ymax * ky, vkz = pi / a.zmax * kz;
          const double vk2 = vkx * vkx + vky * vky + vkz * vkz;
          if (!(vk2 / eta4 < 100.0)) continue;
          const double v = -(4.0 * pi * a.omegain * sin(vkx * (x0 - x1) + vky * (y0 - y1) + vkz * (z0 - z1)) / vk2 *
                             exp(-(vk2 / eta4)));
          fx -= cc * v * vkx;
          fy -= cc * v * vky;
          fz -= cc * v * vkz;
        }
    for (int kz = -a.rsz; kz <= a.rsz; kz++)
      for (int ky = -a.rsy; ky <= a.rsy; ky++)
        for (int kx = -a.rsx; kx <= a.rsx; kx++)
          ewald_real(cc, veta, pi, kx * 2.0 * a.xmax + x0 - x1, ky * 2.0 * a.ymax + y0 - y1,
                     kz * 2.0 * a.zmax + z0 - z1, fx, fy, fz);
  }
  double *o = a.pair + 3 * u;
  o[0] = fx;
  o[1] = fy;
  o[2] = fz;
}

struct FinalArgs {
  int natom, nblk01, nbox_per_atom, nblkp, nblkj, jel;
  const double *p01, *l01, *pcc, *jel_part, *jel_rad, *pair, *cp;
  const int *indspe;
  double biasx, biasy, biasz;
  double *fatx, *faty, *fatz, *fjel;
};

// each atom's terms in the reference's order of groups (force_01, core correction, jellium I,
// core-core), then the bias field; one thread per atom
__global__ void k_force_sum(FinalArgs a) {
  const int na = blockIdx.x * blockDim.x + threadIdx.x;
  if (na >= a.natom) return;
  double fx = a.fatx[na], fy = a.faty[na], fz = a.fatz[na];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int b = 0; b < a.nblk01; b++) {
    const double *o = a.p01 + 3 * ((long)na * a.nblk01 + b);
    sx += o[0];
    sy += o[1];
    sz += o[2];
  }
  sx += a.l01[3 * na];
  sy += a.l01[3 * na + 1];
  sz += a.l01[3 * na + 2];
  fx += sx;
  fy += sy;
  fz += sz;
  double px = 0.0, py = 0.0, pz = 0.0;
  for (long b = 0; b < (long)a.nbox_per_atom * a.nblkp; b++) {
    const double *o = a.pcc + 3 * ((long)na * a.nbox_per_atom * a.nblkp + b);
    px += o[0];
    py += o[1];
    pz += o[2];
  }
  fx += px;
  fy += py;
  fz += pz;
  if (a.jel) {
    double j = 0.0;
    for (int b = 0; b < a.nblkj; b++) j += a.jel_part[(long)na * a.nblkj + b];
    fz += j;  // fatz(na) - fjel0 with fjel0 = -j
  }
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int l1 = 0; l1 < a.natom; l1++) {
    const double *o = a.pair + 3 * ((long)na * a.natom + l1);
    cx += o[0];
    cy += o[1];
    cz += o[2];
  }
  fx += cx;
  fy += cy;
  fz += cz;
  const double c1 = a.cp[8L * (a.indspe[na] - 1)];
  a.fatx[na] = fx + a.biasx * c1;
  a.faty[na] = fy + a.biasy * c1;
  a.fatz[na] = fz + a.biasz * c1;
}

// fjel and the removal of the net force (force8f.F90:1941-2007)
__global__ void k_force_net(FinalArgs a) {
  __shared__ double tot[4];
  if (threadIdx.x == 0) {
    double fj = 0.0;
    if (a.jel) {
      for (int na = 0; na < a.natom; na++) {
        double j = 0.0;
        for (int b = 0; b < a.nblkj; b++) j += a.jel_part[(long)na * a.nblkj + b];
        fj -= j;
      }
      for (int na = 0; na < a.natom; na++) fj += a.jel_rad[na];
      double d = 0.0;
      for (int b = 0; b < a.nblkj; b++) d += a.jel_part[(long)a.natom * a.nblkj + b];
      fj += d;
    }
    double fx = 0.0, fy = 0.0, fz = fj;
    for (int na = 0; na < a.natom; na++) {
      fx += a.fatx[na];
      fy += a.faty[na];
      fz += a.fatz[na];
    }
    tot[0] = fx / (double)a.natom;
    tot[1] = fy / (double)a.natom;
    tot[2] = fz / (double)a.natom;
    tot[3] = fj;
    a.fjel[0] = fj;
  }
  __syncthreads();
  for (int na = threadIdx.x; na < a.natom; na += blockDim.x) {
    a.fatx[na] -= tot[0];
    a.faty[na] -= tot[1];
    a.fatz[na] -= tot[2];
  }
}

}  // namespace

extern "C" {

int kukan_force(int natom, int num_spe, int num_atcell, int num_ppcell_d, int num_list_d, int nqmx, int nradmx,
                int npoint, int nmesh, int nspv, int nperi, int nint1dmax, int nzmax, int jelcalc, int ncpx,
                int ncpy, int ncpz, int ncpx_d, int ncpy_d, int ncpz_d, int new_pwx, int new_pwy, int new_pwz,
                int new_rsx, int new_rsy, int new_rsz, int key_jel_calc, int key_natpri_in, int key_natpri_inps,
                const int *indspe, const int *natpri, const int *natpri_inf, const int *nradct,
                const int *natprid, const int *napsd, const int *natinfd, const int *natinfd_vloc,
                const int *lstvecd2, const int *nqctpcc, const int *lstdx, const int *lstdy, const int *lstdz,
                double veta, double psctoff, double psftrad, const double *radial, const double *dradial,
                const double *point, const double *wt, double xmax, double ymax, double zmax, double biasx,
                double biasy, double biasz, const double *cp, const double *coef, const double *atx,
                const double *aty, const double *atz, const double *dvlocdx_scw, const double *dvlocdy_scw,
                const double *dvlocdz_scw, const double *dvlocdx_hdp, const double *dvlocdy_hdp,
                const double *dvlocdz_hdp, const double *rhotrur, const double *rhosmtr, const double *rhoaugr,
                const double *drhoaug3ddx, const double *drhoaug3ddy, const double *drhoaug3ddz,
                const double *rho_coarse, const double *rho_aug_dense, const double *vloc_dense,
                const double *vh_dense, const double *vx_dense, double chrjel, double strjel, double endjel,
                const int *ndatx, const int *ndaty, const int *ndatz, double *fatx, double *faty,
                double *fatz, double *fjel) {
  API_BEGIN
  if (natom < 1 || num_spe < 1 || nmesh < 1 || nspv < 1 || ncpx < 1 || ncpy < 1 || ncpz < 1 || nradmx < 1)
    throw Error(1, "force: bad sizes");
  if (ncpx_d != ncpx * nmesh || ncpy_d != ncpy * nmesh || ncpz_d != ncpz * nmesh)
    throw Error(1, "force: one rank holds the whole grid (ncpx_d = ncpx * nmesh ...)");
  if (nperi < 0 || nperi > 3) throw Error(1, "force: nperi must be 0..3");
  if (natom > 65534) throw Error(1, "force: more than 65534 atoms");
  if (nperi == 1 && nint1dmax < 1) throw Error(1, "force: nint1dmax must be positive for nperi = 1");
  if (!fjel) throw Error(1, "force: fjel is an output");
  const bool jel = jelcalc == key_jel_calc;
  hipStream_t s = lib().stream;
  const long n = (long)ncpx * ncpy * ncpz, nd = (long)ncpx_d * ncpy_d * ncpz_d;
  const double dx = 2.0 * xmax / (double)ncpx, dy = 2.0 * ymax / (double)ncpy, dz = 2.0 * zmax / (double)ncpz;
  const double ddx = dx / (double)nmesh, ddy = dy / (double)nmesh, ddz = dz / (double)nmesh;
  const double pi = std::acos(-1.0), omega = 8.0 * xmax * ymax * zmax;
  // host copies of the small tables, checked before any device access depends on them
  const std::vector<int> isp = host_ints(indspe, natom), prid = host_ints(natprid, natom),
                         psd = host_ints(napsd, natom), ninfd = host_ints(natinfd, natom),
                         ninfv = host_ints(natinfd_vloc, natom), nrc = host_ints(nradct, num_spe),
                         nqp = host_ints(nqctpcc, num_spe), pri = host_ints(natpri, natom),
                         pinf = host_ints(natpri_inf, natom);
  const std::vector<double> rad = host_doubles(radial, (size_t)nradmx * num_spe), hx = host_doubles(atx, natom),
                            hy = host_doubles(aty, natom), hz = host_doubles(atz, natom);
  for (int na = 0; na < natom; na++) {
    if (isp[na] < 1 || isp[na] > num_spe) throw Error(1, "force: indspe out of range");
    if (prid[na] == key_natpri_inps &&
        (psd[na] < 1 || psd[na] > num_ppcell_d || ninfd[na] < 0 || ninfd[na] > num_list_d || ninfv[na] < 0 ||
         ninfv[na] > num_list_d))
      throw Error(1, "force: napsd / natinfd / natinfd_vloc out of range");
    if (jel && pri[na] == key_natpri_in && (pinf[na] < 1 || pinf[na] > num_atcell))
      throw Error(1, "force: natpri_inf out of range");
  }
  for (int sp = 0; sp < num_spe; sp++) {
    if (nrc[sp] < 1 || nrc[sp] > nradmx) throw Error(1, "force: nradct out of range");
    if (nqp[sp] > nqmx) throw Error(1, "force: nqctpcc exceeds nqmx");
  }
  {
    // list entries index the dense grid
    const std::vector<int> lst = host_ints(lstvecd2, (size_t)num_list_d * num_ppcell_d);
    for (int na = 0; na < natom; na++) {
      if (prid[na] != key_natpri_inps) continue;
      const size_t col = (size_t)num_list_d * (psd[na] - 1);
      for (int p = 0; p < std::max(ninfd[na], ninfv[na]); p++)
        if (lst[col + p] < 1 || lst[col + p] > nd) throw Error(1, "force: lstvecd2 entry outside the dense grid");
    }
  }
  const size_t na3 = (size_t)natom * D, ld = (size_t)num_list_d * num_ppcell_d;
  Arr Ai(indspe, (size_t)natom * I, Arr::In), Apri(natpri, (size_t)natom * I, Arr::In),
      Apinf(natpri_inf, (size_t)natom * I, Arr::In), Anrc(nradct, (size_t)num_spe * I, Arr::In),
      Aprid(natprid, (size_t)natom * I, Arr::In), Apsd(napsd, (size_t)natom * I, Arr::In),
      Ainfd(natinfd, (size_t)natom * I, Arr::In), Ainfv(natinfd_vloc, (size_t)natom * I, Arr::In),
      Alst(lstvecd2, ld * I, Arr::In), Anqp(nqctpcc, (size_t)num_spe * I, Arr::In),
      Alx(lstdx, ld * I, Arr::In), Aly(lstdy, ld * I, Arr::In), Alz(lstdz, ld * I, Arr::In),
      Andx(ndatx, (size_t)natom * I, Arr::In), Andy(ndaty, (size_t)natom * I, Arr::In),
      Andz(ndatz, (size_t)natom * I, Arr::In);
  Arr Arad(radial, (size_t)nradmx * num_spe * D, Arr::In), Adr(dradial, (size_t)nradmx * num_spe * D, Arr::In),
      Apt(point, (size_t)npoint * 3 * D, Arr::In), Awt(wt, (size_t)npoint * D, Arr::In),
      Acp(cp, (size_t)8 * num_spe * D, Arr::In), Acoef(coef, (size_t)(nqmx + 1) * 8 * num_spe * D, Arr::In),
      Ax(atx, na3, Arr::In), Ay(aty, na3, Arr::In), Az(atz, na3, Arr::In);
  Arr Sx(dvlocdx_scw, (size_t)n * natom * D, Arr::In), Sy(dvlocdy_scw, (size_t)n * natom * D, Arr::In),
      Sz(dvlocdz_scw, (size_t)n * natom * D, Arr::In);
  Arr Hx(dvlocdx_hdp, ld * D, Arr::In), Hy(dvlocdy_hdp, ld * D, Arr::In), Hz(dvlocdz_hdp, ld * D, Arr::In);
  Arr Rx(drhoaug3ddx, ld * D, Arr::In), Ry(drhoaug3ddy, ld * D, Arr::In), Rz(drhoaug3ddz, ld * D, Arr::In);
  Arr Rc(rho_coarse, (size_t)n * D, Arr::In), Rad(rho_aug_dense, (size_t)nd * D, Arr::In),
      Vl(vloc_dense, (size_t)nd * D, Arr::In), Vh(vh_dense, (size_t)nd * D, Arr::In),
      Vx(vx_dense, (size_t)nd * nspv * D, Arr::In);
  Arr Fx(fatx, na3, Arr::InOut), Fy(faty, na3, Arr::InOut), Fz(fatz, na3, Arr::InOut);
  const int nblk01 = (int)((n + (long)FT * FPT - 1) / ((long)FT * FPT));
  double *p01 = (double *)scratch("force.p01", (size_t)natom * nblk01 * 3 * D);
  double *l01 = (double *)scratch("force.l01", (size_t)natom * 3 * D);
  {
    Timed t("force_local", 3.0 * 8.0 * (double)n * natom + 8.0 * n);
    hipLaunchKernelGGL(k_f01_soft, dim3(nblk01, natom), dim3(FT), 0, s, n, Sx.d(), Sy.d(), Sz.d(), Rc.d(),
                       dx * dy * dz, p01, nblk01);
    ListF lf{num_list_d, key_natpri_inps, Aprid.i(), Apsd.i(), Ainfd.i(), Ainfv.i(), Alst.i(), Alx.i(), Aly.i(),
             Alz.i(), Andx.i(), Andy.i(), Andz.i(), Hx.d(), Hy.d(), Hz.d(), Rx.d(), Ry.d(), Rz.d(), Rad.d(), Vl.d(),
             Vh.d(), ddx, ddy, ddz, ddx * ddy * ddz, biasx, biasy, biasz, l01};
    hipLaunchKernelGGL(k_f01_lists, dim3(natom), dim3(FT), 0, s, lf);
    HIPCK(hipGetLastError());
  }
  // core correction: the dense-grid box of each atom's 27 image spheres
  std::vector<PccBox> boxes((size_t)natom * 27);
  long maxbox = 0;
  for (int na = 0; na < natom; na++) {
    const int sp = isp[na];
    const double rc = rad[(size_t)(nrc[sp - 1] - 1) + (size_t)nradmx * (sp - 1)];
    const double thr = rc * psctoff, dqq = pi / (rc * psftrad);
    int m = 0;
    for (int kz = -1; kz <= 1; kz++)
      for (int ky = -1; ky <= 1; ky++)
        for (int kx = -1; kx <= 1; kx++, m++) {
          PccBox &b = boxes[(size_t)na * 27 + m];
          b.na = na;
          b.sp = sp;
          b.sx = kx * xmax * 2.0;
          b.sy = ky * ymax * 2.0;
          b.sz = kz * zmax * 2.0;
          b.thr = thr;
          b.dqq = dqq;
          b.vfac = 4.0 * pi * dqq / ((2.0 * pi) * (2.0 * pi) * (2.0 * pi)) / (double)nspv;
          // |i d - (c + h/2 + at - shift)| < thr, widened by one point each side
          auto range = [&](double at, double shift, double cmax, double h, int nmax, int &i0, int &ext) {
            const double c = at - shift + cmax + 0.5 * h;
            const int lo = std::max(1, (int)std::floor((c - thr) / h) - 1);
            const int hi = std::min(nmax, (int)std::ceil((c + thr) / h) + 1);
            i0 = lo;
            ext = std::max(0, hi - lo + 1);
          };
          range(hx[na], b.sx, xmax, ddx, ncpx_d, b.x0, b.bx);
          range(hy[na], b.sy, ymax, ddy, ncpy_d, b.y0, b.by);
          range(hz[na], b.sz, zmax, ddz, ncpz_d, b.z0, b.bz);
          maxbox = std::max(maxbox, (long)b.bx * b.by * b.bz);
        }
  }
  const int nblkp = (int)std::max(1L, std::min(64L, (maxbox + FT - 1) / FT));
  PccBox *boxd = (PccBox *)scratch("force.boxes", boxes.size() * sizeof(PccBox));
  double *pcc = (double *)scratch("force.pcc", boxes.size() * nblkp * 3 * D);
  HIPCK(hipMemcpyAsync(boxd, boxes.data(), boxes.size() * sizeof(PccBox), hipMemcpyHostToDevice, s));
  {
    Timed t("force_pcc", 0.0);
    PccArgs pa{boxd, Acoef.d(), Vx.d(), Ax.d(), Ay.d(), Az.d(), Anqp.i(), nqmx, nspv, ncpx_d, ncpy_d, nd,
               xmax, ymax, zmax, ddx, ddy, ddz, pcc, nblkp};
    hipLaunchKernelGGL(k_pcc, dim3((unsigned)boxes.size(), nblkp), dim3(FT), 0, s, pa);
    HIPCK(hipGetLastError());
  }
  // jellium
  const int nblkj = jel ? (int)std::max(1L, std::min(256L, (nd + FT - 1) / FT)) : 1;
  double *jpart = (double *)scratch("force.jel", (size_t)(natom + 1) * nblkj * D + (size_t)natom * D + ncpz_d * D);
  double *jrad = jpart + (size_t)(natom + 1) * nblkj;
  double *S = jrad + natom;
  if (jel) {
    Timed t("force_jellium", 0.0);
    hipLaunchKernelGGL(k_jel_planes, dim3((ncpz_d + 63) / 64), dim3(64), 0, s, ncpz_d, nzmax, ddz, zmax, pi, omega,
                       chrjel, strjel, endjel, S);
    JelArgs ja{S, Ax.d(), Ay.d(), Az.d(), Acp.d(), Rad.d(), Ai.i(), ncpx_d, ncpy_d, ncpz_d, nd, xmax, ymax, zmax,
               ddx, ddy, ddz, pi, jpart, nblkj};
    hipLaunchKernelGGL(k_jel, dim3(nblkj, natom + 1), dim3(FT), 0, s, ja, natom);
    const size_t rb = (size_t)nradmx * npoint * num_atcell * D;
    Arr Rt(rhotrur, rb * nspv, Arr::In), Rs(rhosmtr, rb * nspv, Arr::In), Ra(rhoaugr, rb, Arr::In);
    JelRad jr{Apri.i(), Apinf.i(), Ai.i(), Anrc.i(), Arad.d(), Adr.d(), Apt.d(), Awt.d(), Az.d(), Rt.d(), Rs.d(),
              Ra.d(), key_natpri_in, nradmx, npoint, nspv, nzmax, zmax, pi, omega, chrjel, strjel, endjel, jrad};
    hipLaunchKernelGGL(k_jel_radial, dim3(natom), dim3(FT), 0, s, jr);
    HIPCK(hipGetLastError());
    sync();  // the radial arrays' device copies go out of scope
  }
  // core-core
  double *pair = (double *)scratch("force.pair", (size_t)natom * natom * 3 * D);
  {
    Timed t("force_core", 0.0);
    CoreArgs ca{Ax.d(), Ay.d(), Az.d(), Acp.d(), Ai.i(), natom, nperi, nint1dmax, new_pwx, new_pwy, new_pwz,
                new_rsx, new_rsy, new_rsz, veta, xmax, ymax, zmax, pi, 1.0 / omega, 4.0 * xmax * ymax, pair};
    const long np = (long)natom * natom;
    hipLaunchKernelGGL(k_core_pairs, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, s, ca);
    HIPCK(hipGetLastError());
  }
  double *dfj = (double *)scratch("force.fjel", D);
  FinalArgs fa{natom, nblk01, 27, nblkp, nblkj, jel ? 1 : 0, p01, l01, pcc, jpart, jrad, pair, Acp.d(), Ai.i(),
               biasx, biasy, biasz, Fx.d(), Fy.d(), Fz.d(), dfj};
  hipLaunchKernelGGL(k_force_sum, dim3((natom + 127) / 128), dim3(128), 0, s, fa);
  hipLaunchKernelGGL(k_force_net, dim3(1), dim3(256), 0, s, fa);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(fjel, dfj, D, is_device_ptr(fjel) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  sync();  // the box table and host copies go out of scope
  Fx.finish();
  Fy.finish();
  Fz.finish();
  API_END
}

}  // extern "C"
