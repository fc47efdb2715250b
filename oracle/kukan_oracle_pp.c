/*
 * kukan_oracle_pp.c — FP64 CPU restatement of the projector and force pipeline of the RSPACE
 * kukan v8f solver: the double-grid weights (pseudocalc_01a), the atom information and list
 * vectors of the projector and compensation-charge spheres (pseudocalc_01b), the
 * double-grid-filtered nonlocal projector table vnlocp (pseudocalc_07) and the local,
 * compensation-charge, core-correction, jellium and core-core forces (force, force_01,
 * force_02).
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like kukan_oracle.c: loaded by the tests, never by
 * the product library.  Every routine follows the reference routine cited above it (file:line
 * relative to RSPACEes.v08.01/kukan/) on one MPI rank (myrx = myry = myrz = 0, one process in
 * each direction) and one OpenMP thread, with its loop and summation order and the operand
 * order of each expression, so that the compiled reference gives the same bits.  That includes
 * the choice of libm entry: where the reference takes the sine and the cosine of one argument
 * the compilers (flang and gcc alike) call glibc's sincos, which differs from sin and cos in the
 * last bit for a small fraction of arguments; where the reference's cosine is dead code (the
 * s-channel projectors of pseudocalc_07) it calls sin alone, and so does the oracle.
 *
 * Pins (tests/test_oracle_pp.py, tests/test_reference_pp.py): the double-grid weights are the
 * Lagrange basis of order 2 nfdg (they reproduce polynomials of degree < 2 nfdg and sum to 1);
 * the list vectors against a numpy brute force of the sphere test; pseudocalc_07 against a
 * numpy evaluation of the Bessel expansion and the separable filter; the core-core forces
 * against finite differences of the Coulomb / Ewald energy and Newton's third law; and every
 * routine against the compiled reference, bit for bit.
 */
#include "kukan_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int orc_set_error(int code, const char *msg); /* kukan_oracle.c */

/* periodic wrap of a 1-based grid index onto 1..n (the do-while loops of pseudocalc8f.F90) */
static int wrap1(int i, int n) {
  while (i > n) i -= n;
  while (i < 1) i += n;
  return i;
}

/* wrap along the periodic axes of nperi (x for nperi >= 1, y for >= 2, z for 3) */
static void wrap3(int nperi, int *i, int *j, int *k, int nx, int ny, int nz) {
  if (nperi >= 1) {
    *i = wrap1(*i, nx);
    if (nperi >= 2) {
      *j = wrap1(*j, ny);
      if (nperi == 3) *k = wrap1(*k, nz);
    }
  }
}

static int inside(int i, int j, int k, int nx, int ny, int nz) {
  return (long)(i - 1) * (nx - i) >= 0 && (long)(j - 1) * (ny - j) >= 0 && (long)(k - 1) * (nz - k) >= 0;
}

/* ------------------------------------------------------------------ pseudocalc8f.F90:343-629
 * Double-grid weights wxyz(j), j = -nfdg*npmesh .. nfdg*npmesh-1 (stored from that lower
 * bound): the Lagrange basis polynomial of the node 0 among the 2 nfdg integer nodes
 * s-nfdg+1 .. s+nfdg around r = |j + 1/2| / npmesh (even npmesh) or |j| / npmesh (odd), s the
 * interval of r (s < r <= s+1, s = 0 for r <= 1).  The product runs over the nodes from the
 * highest down, then divides by the product of (0 - node), as the reference's closed forms
 * do.  For odd npmesh the first element is not used (and is set to zero here). */
int orc_pseudocalc_01a(int nfdg, int npmesh, double *wxyz) {
  if (nfdg < 1 || nfdg > 8) return orc_set_error(1, "error in pseudo. nfdg should be between 1 and 8.");
  if (npmesh < 1) return orc_set_error(1, "pseudocalc_01a: npmesh must be positive");
  const int lo = -nfdg * npmesh, je = nfdg * npmesh - 1;
  int js = lo, ntmp1 = 1;
  double tmp1 = 0.5;
  if (npmesh % 2 == 1) {
    js = lo + 1;
    tmp1 = 0.0;
    ntmp1 = 0;
  }
  const double tmp = 1.0 / npmesh;
  wxyz[0] = 0.0;
  for (int j = js; j <= je; j++) {
    const double r = fabs(j * tmp + tmp1 * tmp);
    int s = 0;
    for (int k = 1; k <= nfdg - 1; k++)
      if (abs(2 * j + ntmp1) > 2 * k * npmesh) s = k;
    double prod = 0.0, den = 1.0;
    int first = 1;
    for (int m = s + nfdg; m >= s - nfdg + 1; m--) {
      if (m == 0) continue;
      const double f = r - (double)m;
      prod = first ? f : prod * f;
      first = 0;
      den *= (double)(-m);
    }
    wxyz[j - lo] = prod / den;
  }
  return 0;
}

/* ------------------------------------------------------------------ pseudocalc8f.F90:633-1159
 * Atom information and list vectors on one rank.  natx.. = the coarse point nearest the atom
 * (relative to the cell centre), ndatx.. the same on the dense grid; natpri = in for atoms
 * whose point lies in the (single) subdomain, else inps when a point of the projector sphere
 * (test of the (2 nfdg + 1)^3 neighbourhood against radial(nradct) * psctoff, :811-826) maps
 * into the grid, else out; natprid = inps when a dense point of the sphere maps into the
 * dense grid.  lstvec2 / lstx.. list the coarse points of each in/inps atom's box whose
 * neighbourhood meets the sphere, in z-y-x order with x fastest (:954-1033); lstvecd2 /
 * lstdx.. list the dense points inside radial(nradct) (natinfd of them) followed by those in
 * the shell up to radial(nradct) * psctoff (natinfd_vloc in all, :1036-1150).  Columns of
 * atoms that are not listed are left as they are. */
int orc_pseudocalc_01b(int nperi, int natom, int num_spe, int nradmx, int nmesh, int nfdg,
                       int ndisp, int num_atcell, int num_ppcell, int num_ppcell_d, int num_list,
                       int num_list_d, int nxmax, int nymax, int nzmax, int ncpx, int ncpy,
                       int ncpz, int npxmax, int npymax, int npzmax, int key_natpri_in,
                       int key_natpri_inps, int key_natpri_out, double psctoff, double xmax,
                       double ymax, double zmax, double dx, double dy, double dz,
                       const int *indspe, const int *nradct, const double *radial,
                       const double *atx, const double *aty, const double *atz, int *natx,
                       int *naty, int *natz, int *ndatx, int *ndaty, int *ndatz, int *lstx,
                       int *lsty, int *lstz, int *lstdx, int *lstdy, int *lstdz, int *natinf,
                       int *natinfd, int *natinfd_vloc, int *lstvec2, int *lstvecd2, int *natpri,
                       int *natprid, int *natpri_inf, int *naps, int *napsd) {
  (void)ndisp;
  (void)num_spe;
  const int ncpx_d = ncpx * nmesh, ncpy_d = ncpy * nmesh, ncpz_d = ncpz * nmesh;
  const int npxmax_d = npxmax * nmesh, npymax_d = npymax * nmesh, npzmax_d = npzmax * nmesh;
  const double ddx = dx / nmesh, ddy = dy / nmesh, ddz = dz / nmesh;
  int l_atcell = 1, l_ppcell = 1, l_ppcell_d = 1;
#define RC(na) radial[(size_t)(nradct[indspe[na] - 1] - 1) + (size_t)nradmx * (indspe[na] - 1)]

  for (int na = 0; na < natom; na++) {
    natx[na] = (int)((atx[na] + xmax + 0.5 * dx) / dx) - nxmax;
    naty[na] = (int)((aty[na] + ymax + 0.5 * dy) / dy) - nymax;
    natz[na] = (int)((atz[na] + zmax + 0.5 * dz) / dz) - nzmax;
    ndatx[na] = (int)((atx[na] + xmax + 0.5 * ddx) / ddx) - nxmax * nmesh;
    ndaty[na] = (int)((aty[na] + ymax + 0.5 * ddy) / ddy) - nymax * nmesh;
    ndatz[na] = (int)((atz[na] + zmax + 0.5 * ddz) / ddz) - nzmax * nmesh;
  }
  for (int na = 0; na < natom; na++) {
    const int i = (natx[na] + nxmax - 1) / ncpx, j = (naty[na] + nymax - 1) / ncpy,
              k = (natz[na] + nzmax - 1) / ncpz;
    natpri[na] = (i == 0 && j == 0 && k == 0) ? key_natpri_in : key_natpri_out;
  }
  {
    int l = 0;
    for (int na = 0; na < natom; na++) {
      natpri_inf[na] = 0;
      if (natpri[na] == key_natpri_in) {
        natpri_inf[na] = ++l;
        if (natpri_inf[na] > num_atcell) l_atcell = 0;
      }
    }
  }
  /* projector sphere of an out atom reaching into the grid (:789-883, the one-rank branch) */
  for (int na = 0; na < natom; na++) {
    const double ax = atx[na] - (natx[na] * dx - 0.5 * dx), ay = aty[na] - (naty[na] * dy - 0.5 * dy),
                 az = atz[na] - (natz[na] * dz - 0.5 * dz);
    const double thr = RC(na) * psctoff;
    int l0 = 0;
    for (int iz = -npzmax + 1; iz <= npzmax; iz++)
      for (int iy = -npymax + 1; iy <= npymax; iy++)
        for (int ix = -npxmax + 1; ix <= npxmax; ix++) {
          if (natpri[na] != key_natpri_out) continue;
          int l = 0;
          for (int jz = -nfdg; jz <= nfdg; jz++)
            for (int jy = -nfdg; jy <= nfdg; jy++)
              for (int jx = -nfdg; jx <= nfdg; jx++) {
                const double x = (ix + jx) * dx - ax, y = (iy + jy) * dy - ay, z = (iz + jz) * dz - az;
                if (sqrt(x * x + y * y + z * z) < thr) l = 1;
              }
          if (l) {
            int i = ix + natx[na] + nxmax, j = iy + naty[na] + nymax, k = iz + natz[na] + nzmax;
            wrap3(nperi, &i, &j, &k, 2 * nxmax, 2 * nymax, 2 * nzmax);
            if (inside(i, j, k, ncpx, ncpy, ncpz)) l0 = 1;
          }
        }
    if (natpri[na] == key_natpri_out && l0) natpri[na] = key_natpri_inps;
  }
  {
    int iaps = 0;
    for (int na = 0; na < natom; na++) {
      naps[na] = 0;
      if (natpri[na] == key_natpri_in || natpri[na] == key_natpri_inps) {
        naps[na] = ++iaps;
        if (naps[na] > num_ppcell) l_ppcell = 0;
      }
    }
  }
  /* dense sphere in the grid (:889-933) */
  for (int na = 0; na < natom; na++) {
    natprid[na] = key_natpri_out;
    const double thr = RC(na) * psctoff;
    int l0 = 0;
    for (int iz = -npzmax_d + 1; iz <= npzmax_d; iz++)
      for (int iy = -npymax_d + 1; iy <= npymax_d; iy++)
        for (int ix = -npxmax_d + 1; ix <= npxmax_d; ix++) {
          int i = ix + ndatx[na] + nxmax * nmesh, j = iy + ndaty[na] + nymax * nmesh,
              k = iz + ndatz[na] + nzmax * nmesh;
          const double x = ix * ddx - (atx[na] - (ndatx[na] * ddx - 0.5 * ddx));
          const double y = iy * ddy - (aty[na] - (ndaty[na] * ddy - 0.5 * ddy));
          const double z = iz * ddz - (atz[na] - (ndatz[na] * ddz - 0.5 * ddz));
          if (sqrt(x * x + y * y + z * z) < thr) {
            wrap3(nperi, &i, &j, &k, 2 * nxmax * nmesh, 2 * nymax * nmesh, 2 * nzmax * nmesh);
            if (inside(i, j, k, ncpx_d, ncpy_d, ncpz_d)) l0 = 1;
          }
        }
    if (l0) natprid[na] = key_natpri_inps;
  }
  {
    int iapsd = 0;
    for (int na = 0; na < natom; na++) {
      napsd[na] = 0;
      if (natprid[na] == key_natpri_inps) {
        napsd[na] = ++iapsd;
        if (napsd[na] > num_ppcell_d) l_ppcell_d = 0;
      }
    }
  }
  if (!l_atcell) return orc_set_error(1, "pseudocalc_01b: num_atcell is too small.");
  if (!l_ppcell) return orc_set_error(1, "error! num_ppcell is too small.");
  if (!l_ppcell_d) return orc_set_error(1, "error! num_ppcell_d is too small.");

  /* projector list vectors (:954-1033) */
  for (int na = 0; na < natom; na++) {
    if (natpri[na] != key_natpri_in && natpri[na] != key_natpri_inps) {
      natinf[na] = 0;
      continue;
    }
    const size_t col = (size_t)num_list * (naps[na] - 1);
    const double ax = atx[na] - (natx[na] * dx - 0.5 * dx), ay = aty[na] - (naty[na] * dy - 0.5 * dy),
                 az = atz[na] - (natz[na] * dz - 0.5 * dz);
    const double thr = RC(na) * psctoff;
    for (int l = 0; l < num_list; l++) lstvec2[col + l] = lstx[col + l] = lsty[col + l] = lstz[col + l] = 0;
    int l0 = 0;
    for (int iz = -npzmax + 1; iz <= npzmax; iz++)
      for (int iy = -npymax + 1; iy <= npymax; iy++)
        for (int ix = -npxmax + 1; ix <= npxmax; ix++) {
          int l = 0;
          for (int jz = -nfdg; jz <= nfdg; jz++)
            for (int jy = -nfdg; jy <= nfdg; jy++)
              for (int jx = -nfdg; jx <= nfdg; jx++) {
                const double x = (ix + jx) * dx - ax, y = (iy + jy) * dy - ay, z = (iz + jz) * dz - az;
                if (sqrt(x * x + y * y + z * z) < thr) l = 1;
              }
          if (!l) continue;
          int i = ix + natx[na] + nxmax, j = iy + naty[na] + nymax, k = iz + natz[na] + nzmax;
          wrap3(nperi, &i, &j, &k, 2 * nxmax, 2 * nymax, 2 * nzmax);
          if (!inside(i, j, k, ncpx, ncpy, ncpz)) continue;
          l0++;
          if (l0 <= num_list) {
            lstvec2[col + l0 - 1] = (k - 1) * ncpy * ncpx + (j - 1) * ncpx + i;
            lstx[col + l0 - 1] = ix;
            lsty[col + l0 - 1] = iy;
            lstz[col + l0 - 1] = iz;
          }
        }
    natinf[na] = l0;
    if (natinf[na] > num_list) return orc_set_error(1, "error! num_list is too small.");
  }

  /* compensation-charge / hard local list vectors on the dense grid (:1036-1150) */
  for (int na = 0; na < natom; na++) {
    if (natprid[na] != key_natpri_inps) {
      natinfd[na] = 0;
      natinfd_vloc[na] = 0;
      continue;
    }
    const size_t col = (size_t)num_list_d * (napsd[na] - 1);
    const double rc = RC(na), thr = rc * psctoff;
    for (int l = 0; l < num_list_d; l++)
      lstvecd2[col + l] = lstdx[col + l] = lstdy[col + l] = lstdz[col + l] = 0;
    int l0 = 0;
    for (int pass = 0; pass < 2; pass++) {
      for (int iz = -npzmax_d + 1; iz <= npzmax_d; iz++)
        for (int iy = -npymax_d + 1; iy <= npymax_d; iy++)
          for (int ix = -npxmax_d + 1; ix <= npxmax_d; ix++) {
            int i = ix + ndatx[na] + nxmax * nmesh, j = iy + ndaty[na] + nymax * nmesh,
                k = iz + ndatz[na] + nzmax * nmesh;
            const double x = ix * ddx - (atx[na] - (ndatx[na] * ddx - 0.5 * ddx));
            const double y = iy * ddy - (aty[na] - (ndaty[na] * ddy - 0.5 * ddy));
            const double z = iz * ddz - (atz[na] - (ndatz[na] * ddz - 0.5 * ddz));
            const double r = sqrt(x * x + y * y + z * z);
            if (pass == 0 ? !(r < rc) : !(r >= rc && r < thr)) continue;
            wrap3(nperi, &i, &j, &k, 2 * nxmax * nmesh, 2 * nymax * nmesh, 2 * nzmax * nmesh);
            if (!inside(i, j, k, ncpx_d, ncpy_d, ncpz_d)) continue;
            l0++;
            if (l0 <= num_list_d) {
              lstvecd2[col + l0 - 1] = (k - 1) * ncpy_d * ncpx_d + (j - 1) * ncpx_d + i;
              lstdx[col + l0 - 1] = ix;
              lstdy[col + l0 - 1] = iy;
              lstdz[col + l0 - 1] = iz;
            }
          }
      if (pass == 0) natinfd[na] = l0;
    }
    natinfd_vloc[na] = l0;
    if (natinfd_vloc[na] > num_list_d) return orc_set_error(1, "error! num_list_d is too small.");
  }
#undef RC
  return 0;
}

/* Fermi filter of mathfunctions8f.F90:21-36 with weight 1 */
static double fermidis1(double r, double t, double rc) {
  const double a = (r - rc) / t;
  double f;
  if (fabs(a) < 1.0e2) f = 1.0 / (exp(a) + 1.0);
  else f = a > 0.0 ? 0.0 : 1.0;
  return f * 1.0;
}

/* the radial channels (coef index 1..6: two l=0, two l=1, two l=2) present for each nprj
 * (the cases of pseudocalc8f.F90:2589-3620; other nprj values have no case) */
static int pp_channels(int nprj, int ch[6]) {
  static const struct { int nprj, mask; } T[] = {
      {1, 0x01}, {2, 0x03}, {4, 0x05}, {5, 0x07}, {7, 0x0d}, {8, 0x0f}, {9, 0x15},
      {10, 0x17}, {12, 0x1d}, {13, 0x1f}, {14, 0x35}, {15, 0x37}, {17, 0x3d}, {18, 0x3f}};
  for (size_t t = 0; t < sizeof T / sizeof T[0]; t++)
    if (T[t].nprj == nprj) {
      for (int c = 0; c < 6; c++) ch[c] = (T[t].mask >> c) & 1;
      return 1;
    }
  return 0;
}

/* ------------------------------------------------------------------ pseudocalc8f.F90:2515-3702
 * vnlocp(:, la, naps(na)) for every in/inps atom: the projectors (Bessel expansion of each
 * radial channel, coef(iq, c, spe) q^2 j_l(q r), q = iq pi / rcutss, times the real
 * harmonic of each m and the optional Fermi filter) on the fine grid of spacing dx/npmesh
 * around the atom (vnlspd), then the double-grid filter of weights wxyz along z, y and x onto
 * the coarse points, read out at the atom's list points.  vtmp_pp, vtmp_pp1, vtmp_pp2 and
 * vnlspd are the caller's work arrays, as in the reference. */
int orc_pseudocalc_07(int natom, int num_spe, int npmesh, int nprjmx, int nfdg, int nfiltyp,
                      int num_list, int num_ppcell, int nradmx, int nqmx, int nxmax, int nymax,
                      int nzmax, int npxmax, int npymax, int npzmax, int nxspd, int nyspd,
                      int nzspd, int key_natpri_in, int key_natpri_inps, double eps,
                      double psftrad, double psctoff, double filpp, double dx, double dy,
                      double dz, const int *indspe, const int *nprj, const int *nradct,
                      const int *natpri, const int *naps, const int *natinf, const int *nqct,
                      const int *natx, const int *naty, const int *natz, const int *lstx,
                      const int *lsty, const int *lstz, const double *wxyz, const double *radial,
                      const double *coef, const double *atx, const double *aty,
                      const double *atz, double *vnlocp, double *vtmp_pp, double *vtmp_pp1,
                      double *vtmp_pp2, double *vnlspd) {
  (void)num_spe;
  const double pi = acos(-1.0);
  const double sqfourpi = sqrt(1.0 / 4.0 / pi), sq3fourpi = sqrt(3.0 / 4.0 / pi),
               sq1516pi = sqrt(15.0 / 16.0 / pi), sq1504pi = sqrt(15.0 / 4.0 / pi),
               sq0516pi = sqrt(5.0 / 16.0 / pi);
  const double dpx = dx / npmesh, dpy = dy / npmesh, dpz = dz / npmesh;
  const double dpxyz = dpx * dpy * dpz;
  const int npxmax_d = npxmax * npmesh, npymax_d = npymax * npmesh, npzmax_d = npzmax * npmesh;
  const double twopicbin = 1.0 / ((2.0 * pi) * (2.0 * pi) * (2.0 * pi));
  const double fourpi = 4.0 * pi;
  /* array extents */
  const long sx = 2L * nxspd, sy = 2L * nyspd, sz = 2L * nzspd;
  const long cx = 2L * npxmax, cy = 2L * npymax, cz = 2L * npzmax;
  const int wlo = -nfdg * npmesh;
#define VSPD(kx, ky, kz, la) vnlspd[((kx) + nxspd - 1) + sx * (((ky) + nyspd - 1) + sy * (((kz) + nzspd - 1) + sz * (long)((la) - 1)))]
#define VPP(ix, iy, iz) vtmp_pp[((ix) + npxmax - 1) + cx * (((iy) + npymax - 1) + cy * (long)((iz) + npzmax - 1))]
#define VPP1(ix, iy, iz) vtmp_pp1[((ix) + nxspd - 1) + sx * (((iy) + nyspd - 1) + sy * (long)((iz) + npzmax - 1))]
#define VPP2(ix, iy, iz) vtmp_pp2[((ix) + nxspd - 1) + sx * (((iy) + npymax - 1) + cy * (long)((iz) + npzmax - 1))]
#define W(j) wxyz[(j) - wlo]
#define COEF(iq, c, sp) coef[(size_t)(iq) + (size_t)(nqmx + 1) * ((c) + 8 * (size_t)((sp) - 1))]
  memset(vnlocp, 0, sizeof(double) * (size_t)num_list * nprjmx * num_ppcell);
  for (int na = 0; na < natom; na++) {
    memset(vnlspd, 0, sizeof(double) * (size_t)sx * sy * sz * nprjmx);
    if (natpri[na] != key_natpri_in && natpri[na] != key_natpri_inps) continue;
    const int sp = indspe[na];
    const int npatx = (int)((atx[na] + nxmax * dx + 0.5 * dpx) / dpx) - nxmax * npmesh;
    const int npaty = (int)((aty[na] + nymax * dy + 0.5 * dpy) / dpy) - nymax * npmesh;
    const int npatz = (int)((atz[na] + nzmax * dz + 0.5 * dpz) / dpz) - nzmax * npmesh;
    const int iaps = naps[na];
    const double rc = radial[(size_t)(nradct[sp - 1] - 1) + (size_t)nradmx * (sp - 1)];
    const double rcutss = rc * psftrad;
    const int np = nprj[sp - 1];
    int ch[6];
    if (pp_channels(np, ch)) {
      const double dqq = pi / rcutss;
      const double tmp = fourpi * dqq * twopicbin;
      const int nq = nqct[sp - 1];
      for (int iz = -npzmax_d + 1; iz <= npzmax_d; iz++)
        for (int iy = -npymax_d + 1; iy <= npymax_d; iy++)
          for (int ix = -npxmax_d + 1; ix <= npxmax_d; ix++) {
            const int kx = ix + (npatx - npmesh * natx[na]);
            const int ky = iy + (npaty - npmesh * naty[na]);
            const int kz = iz + (npatz - npmesh * natz[na]);
            const double x = ix * dpx - (atx[na] - (npatx * dpx - 0.5 * dpx));
            const double y = iy * dpy - (aty[na] - (npaty * dpy - 0.5 * dpy));
            const double z = iz * dpz - (atz[na] - (npatz * dpz - 0.5 * dpz));
            const double r = sqrt(x * x + y * y + z * z);
            if (!(r < rc * psctoff)) continue;
            double t[7] = {0, 0, 0, 0, 0, 0, 0};
            double ylm[10] = {0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0};
            double weight = 1.0;
            if (r > eps) {
              if (nfiltyp == 1) weight = fermidis1(r, filpp, rc);
              const double rin = 1.0 / r;
              ylm[2] = rin * x;
              ylm[3] = rin * y;
              ylm[4] = rin * z;
              ylm[5] = rin * rin * (x * x - y * y);
              ylm[6] = -(rin * rin * z * x);
              ylm[7] = rin * rin * (3.0 * z * z - r * r);
              ylm[8] = -(rin * rin * y * z);
              ylm[9] = rin * rin * x * y;
              if (!(ch[2] || ch[3] || ch[4] || ch[5])) {
                /* s channels only: the reference's cos is dead code there, so it calls sin
                 * alone; glibc's sincos (what compilers make of a sin and a cos of one argument)
                 * differs from sin in the last bit for some arguments */
                for (int iq = 1; iq <= nq; iq++) {
                  const double qqq = dqq * iq;
                  const double qqq2 = qqq * qqq;
                  const double qqqr = qqq * r;
                  const double qqqrin = rin / qqq;
                  const double j0 = sin(qqqr) * qqqrin;
                  if (np == 1) {
                    /* the nprj = 1 case multiplies by q twice (pseudocalc8f.F90:2615) */
                    t[1] = t[1] + COEF(iq, 1, sp) * j0 * qqq * qqq;
                  } else {
                    if (ch[0]) t[1] = t[1] + COEF(iq, 1, sp) * j0 * qqq2;
                    if (ch[1]) t[2] = t[2] + COEF(iq, 2, sp) * j0 * qqq2;
                  }
                }
              } else {
                for (int iq = 1; iq <= nq; iq++) {
                  const double qqq = dqq * iq;
                  const double qqq2 = qqq * qqq;
                  const double qqqr = qqq * r;
                  const double qqqr2 = qqqr * qqqr;
                  const double qqqrin = rin / qqq;
                  const double qqqr2in = qqqrin * qqqrin;
                  const double qqqr3in = qqqr2in * qqqrin;
                  const double s = sin(qqqr), c = cos(qqqr);
                  const double j0 = s * qqqrin;
                  const double j1 = (s - qqqr * c) * qqqr2in;
                  const double j2 = ((3.0 - qqqr2) * s - 3.0 * qqqr * c) * qqqr3in;
                  if (ch[0]) t[1] = t[1] + COEF(iq, 1, sp) * j0 * qqq2;
                  if (ch[1]) t[2] = t[2] + COEF(iq, 2, sp) * j0 * qqq2;
                  if (ch[2]) t[3] = t[3] + COEF(iq, 3, sp) * j1 * qqq2;
                  if (ch[3]) t[4] = t[4] + COEF(iq, 4, sp) * j1 * qqq2;
                  if (ch[4]) t[5] = t[5] + COEF(iq, 5, sp) * j2 * qqq2;
                  if (ch[5]) t[6] = t[6] + COEF(iq, 6, sp) * j2 * qqq2;
                }
              }
            } else {
              for (int iq = 1; iq <= nq; iq++) {
                const double qqq = dqq * iq;
                if (ch[0]) t[1] = t[1] + COEF(iq, 1, sp) * qqq * qqq;
                if (ch[1]) t[2] = t[2] + COEF(iq, 2, sp) * qqq * qqq;
              }
            }
            /* projector order: channels in order, each with its 2l+1 harmonics */
            int la = 1;
            for (int c = 1; c <= 6; c++) {
              if (!ch[c - 1]) continue;
              const int l = (c - 1) / 2;
              if (l == 0) {
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sqfourpi * ylm[1] * weight;
              } else if (l == 1) {
                for (int m = 2; m <= 4; m++) VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq3fourpi * ylm[m] * weight;
              } else {
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq1516pi * ylm[5] * weight;
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq1504pi * ylm[6] * weight;
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq0516pi * ylm[7] * weight;
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq1504pi * ylm[8] * weight;
                VSPD(kx, ky, kz, la++) = t[c] * tmp * dpxyz * sq1504pi * ylm[9] * weight;
              }
            }
          }
    }
    /* the double-grid filter, z then y then x (:3662-3697) */
    const int je = nfdg * npmesh - 1;
    const int js = (npmesh % 2 == 1) ? -nfdg * npmesh + 1 : -nfdg * npmesh;
    const int half = (npmesh - 1) / 2;
    for (int la = 1; la <= np; la++) {
      memset(vtmp_pp, 0, sizeof(double) * (size_t)cx * cy * cz);
      memset(vtmp_pp1, 0, sizeof(double) * (size_t)sx * sy * cz);
      memset(vtmp_pp2, 0, sizeof(double) * (size_t)sx * cy * cz);
      for (int iz = -npzmax + 1; iz <= npzmax; iz++)
        for (int iy = -nyspd + 1; iy <= nyspd; iy++)
          for (int ix = -nxspd + 1; ix <= nxspd; ix++) {
            const int kkz = iz * npmesh - half;
            for (int jz = js; jz <= je; jz++) VPP1(ix, iy, iz) = VPP1(ix, iy, iz) + VSPD(ix, iy, kkz + jz, la) * W(jz);
          }
      for (int iz = -npzmax + 1; iz <= npzmax; iz++)
        for (int iy = -npymax + 1; iy <= npymax; iy++)
          for (int ix = -nxspd + 1; ix <= nxspd; ix++) {
            const int kky = iy * npmesh - half;
            for (int jy = js; jy <= je; jy++) VPP2(ix, iy, iz) = VPP2(ix, iy, iz) + VPP1(ix, kky + jy, iz) * W(jy);
          }
      for (int iz = -npzmax + 1; iz <= npzmax; iz++)
        for (int iy = -npymax + 1; iy <= npymax; iy++)
          for (int ix = -npxmax + 1; ix <= npxmax; ix++) {
            const int kkx = ix * npmesh - half;
            for (int jx = js; jx <= je; jx++) VPP(ix, iy, iz) = VPP(ix, iy, iz) + VPP2(kkx + jx, iy, iz) * W(jx);
          }
      const size_t col = (size_t)num_list * ((la - 1) + (size_t)nprjmx * (iaps - 1));
      const size_t lcol = (size_t)num_list * (iaps - 1);
      for (int p = 0; p < natinf[na]; p++)
        vnlocp[col + p] = VPP(lstx[lcol + p], lsty[lcol + p], lstz[lcol + p]);
    }
  }
#undef VSPD
#undef VPP
#undef VPP1
#undef VPP2
#undef W
#undef COEF
  return 0;
}

/* ------------------------------------------------------------------ force8f.F90:2016-2109
 * The local and compensation-charge parts of the force on atom na (one thread): the soft
 * local potential's derivative against the coarse density, its hard part against the dense
 * augmented density on the atom's vloc list, and the compensation charge's derivative
 * against vloc + v_H + the bias field on the atom's sphere list. */
static void force_01(int na, int nmesh, int ncpx, int ncpy, int ncpz, int num_list_d,
                     int key_natpri_inps, double dx, double dy, double dz, double *tx, double *ty,
                     double *tz, const int *natprid, const int *napsd, const int *natinfd,
                     const int *natinfd_vloc, const int *lstvecd2, const int *lstdx,
                     const int *lstdy, const int *lstdz, const double *rho_coarse,
                     const double *rho_aug_dense, const double *vh_dense, const double *drx,
                     const double *dry, const double *drz, const double *dvx_scw,
                     const double *dvy_scw, const double *dvz_scw, const double *dvx_hdp,
                     const double *dvy_hdp, const double *dvz_hdp, const int *ndatx,
                     const int *ndaty, const int *ndatz, double biasx, double biasy, double biasz,
                     const double *vloc_dense) {
  const double ddx = dx / (double)nmesh, ddy = dy / (double)nmesh, ddz = dz / (double)nmesh;
  double x0 = 0.0, y0 = 0.0, z0 = 0.0;
  const double dxyz = dx * dy * dz;
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  const double *ax = dvx_scw + n * na, *ay = dvy_scw + n * na, *az = dvz_scw + n * na;
  for (size_t p = 0; p < n; p++) {
    x0 = x0 + ax[p] * rho_coarse[p] * dxyz;
    y0 = y0 + ay[p] * rho_coarse[p] * dxyz;
    z0 = z0 + az[p] * rho_coarse[p] * dxyz;
  }
  if (natprid[na] == key_natpri_inps) {
    const double ddxyz = ddx * ddy * ddz;
    const size_t col = (size_t)num_list_d * (napsd[na] - 1);
    for (int p = 0; p < natinfd_vloc[na]; p++) {
      const size_t i = (size_t)lstvecd2[col + p] - 1;
      x0 = x0 + dvx_hdp[col + p] * rho_aug_dense[i] * ddxyz;
      y0 = y0 + dvy_hdp[col + p] * rho_aug_dense[i] * ddxyz;
      z0 = z0 + dvz_hdp[col + p] * rho_aug_dense[i] * ddxyz;
    }
    for (int p = 0; p < natinfd[na]; p++) {
      const size_t i = (size_t)lstvecd2[col + p] - 1;
      const double x = lstdx[col + p] * ddx + (ndatx[na] * ddx - 0.5 * ddx);
      const double y = lstdy[col + p] * ddy + (ndaty[na] * ddy - 0.5 * ddy);
      const double z = lstdz[col + p] * ddz + (ndatz[na] * ddz - 0.5 * ddz);
      const double v = vloc_dense[i] + vh_dense[i] + biasx * x + biasy * y + biasz * z;
      x0 = x0 + v * drx[col + p] * ddxyz;
      y0 = y0 + v * dry[col + p] * ddxyz;
      z0 = z0 + v * drz[col + p] * ddxyz;
    }
  }
  *tx = *tx + x0;
  *ty = *ty + y0;
  *tz = *tz + z0;
}

/* ------------------------------------------------------------------ force8f.F90:2112-2190
 * Partial-core-correction force: the radial derivative of the core density (Bessel expansion
 * of coef(:, 7, spe), nqctpcc terms) against v_xc on the dense grid, over the 27 nearest
 * periodic images of each atom and each spin. */
static void force_02(int natom, int nspv, int nradmx, int nqmx, int ncpx_d, int ncpy_d,
                     int ncpz_d, double xmax, double ymax, double zmax, double ddx, double ddy,
                     double ddz, double psctoff, double psftrad, const int *nradct,
                     const int *indspe, const int *nqctpcc, double *fatx, double *faty,
                     double *fatz, const double *atx, const double *aty, const double *atz,
                     const double *radial, const double *coef, const double *vx_dense) {
  const double pi = acos(-1.0);
  const double twopicbin = 1.0 / ((2.0 * pi) * (2.0 * pi) * (2.0 * pi));
  const double fourpi = 4.0 * pi;
  const size_t nd = (size_t)ncpx_d * ncpy_d * ncpz_d;
  for (int na = 0; na < natom; na++) {
    const int sp = indspe[na];
    const double rc = radial[(size_t)(nradct[sp - 1] - 1) + (size_t)nradmx * (sp - 1)];
    const double rcutss = rc * psftrad;
    for (int ns = 0; ns < nspv; ns++)
      for (int kz = -1; kz <= 1; kz++)
        for (int ky = -1; ky <= 1; ky++)
          for (int kx = -1; kx <= 1; kx++) {
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int iz = 1; iz <= ncpz_d; iz++)
              for (int iy = 1; iy <= ncpy_d; iy++)
                for (int ix = 1; ix <= ncpx_d; ix++) {
                  const double x = ix * ddx - xmax - 0.5 * ddx - atx[na] + kx * xmax * 2.0;
                  const double y = iy * ddy - ymax - 0.5 * ddy - aty[na] + ky * ymax * 2.0;
                  const double z = iz * ddz - zmax - 0.5 * ddz - atz[na] + kz * zmax * 2.0;
                  const double r = sqrt(x * x + y * y + z * z);
                  double tmp1 = 0.0;
                  if (r < rc * psctoff) {
                    const double dqq = pi / rcutss;
                    const double rin = 1.0 / r;
                    for (int iq = 1; iq <= nqctpcc[sp - 1]; iq++) {
                      const double qqq = dqq * iq;
                      const double qqqr = qqq * r;
                      const double qqqrin = rin / qqq;
                      const double db = -(sin(qqqr) * qqqrin * rin) + cos(qqqr) * rin;
                      tmp1 = tmp1 + coef[(size_t)iq + (size_t)(nqmx + 1) * (7 + 8 * (size_t)(sp - 1))] * db * qqq * qqq;
                    }
                    const double vlo = tmp1 * fourpi * dqq * twopicbin / (double)nspv;
                    const double vcc = vx_dense[(size_t)(ix - 1) + (size_t)ncpx_d * ((iy - 1) + (size_t)ncpy_d * (iz - 1)) + nd * ns];
                    fx = fx + vlo * vcc * ddx * ddy * ddz * (x / r);
                    fy = fy + vlo * vcc * ddx * ddy * ddz * (y / r);
                    fz = fz + vlo * vcc * ddx * ddy * ddz * (z / r);
                  }
                }
            fatx[na] = fatx[na] + fx;
            faty[na] = faty[na] + fy;
            fatz[na] = fatz[na] + fz;
          }
  }
}

/* the real-space Ewald term of one image (force8f.F90:1843-1851, 1874-1882, 1909-1921) */
static void ewald_real(double cc, double veta, double pi, double x, double y, double z,
                       double *fx, double *fy, double *fz) {
  const double r = sqrt(x * x + y * y + z * z);
  if (r > 1.0e-20 && (veta * veta) * (r * r) < 100.0) {
    const double vlo0 = -(2.0 * veta / (exp((veta * veta) * (r * r)) * sqrt(pi) * r)) - (1.0 - erf(veta * r)) / (r * r);
    *fx = *fx - cc * vlo0 * x / r;
    *fy = *fy - cc * vlo0 * y / r;
    *fz = *fz - cc * vlo0 * z / r;
  }
}

/* ------------------------------------------------------------------ force8f.F90:1617-2013
 * Forces on the atoms beyond the band-structure term (which force_eig_r adds to fatx..):
 * force_01 per atom, the core-correction force, the jellium slab (jelcalc = key_jel_calc),
 * the core-core interaction (Coulomb for nperi = 0; Ewald sums in 1, 2 and 3 periodic
 * directions with new_pw* reciprocal and new_rs* real-space images), the bias field, and
 * the removal of the net force. */
int orc_force(int natom, int num_spe, int num_atcell, int num_ppcell_d, int num_list_d, int nqmx,
              int nradmx, int npoint, int nmesh, int nspv, int nperi, int nint1dmax, int nzmax,
              int jelcalc, int ncpx, int ncpy, int ncpz, int ncpx_d, int ncpy_d, int ncpz_d,
              int new_pwx, int new_pwy, int new_pwz, int new_rsx, int new_rsy, int new_rsz,
              int key_jel_calc, int key_natpri_in, int key_natpri_inps, const int *indspe,
              const int *natpri, const int *natpri_inf, const int *nradct, const int *natprid,
              const int *napsd, const int *natinfd, const int *natinfd_vloc,
              const int *lstvecd2, const int *nqctpcc, const int *lstdx, const int *lstdy,
              const int *lstdz, double veta, double psctoff, double psftrad,
              const double *radial, const double *dradial, const double *point,
              const double *wt, double xmax, double ymax, double zmax, double biasx,
              double biasy, double biasz, const double *cp, const double *coef,
              const double *atx, const double *aty, const double *atz,
              const double *dvlocdx_scw, const double *dvlocdy_scw, const double *dvlocdz_scw,
              const double *dvlocdx_hdp, const double *dvlocdy_hdp, const double *dvlocdz_hdp,
              const double *rhotrur, const double *rhosmtr, const double *rhoaugr,
              const double *drhoaug3ddx, const double *drhoaug3ddy, const double *drhoaug3ddz,
              const double *rho_coarse, const double *rho_aug_dense, const double *vloc_dense,
              const double *vh_dense, const double *vx_dense, double chrjel, double strjel,
              double endjel, const int *ndatx, const int *ndaty, const int *ndatz,
              double *fatx, double *faty, double *fatz, double *fjel) {
  (void)num_spe;
  (void)num_atcell;
  (void)num_ppcell_d;
  const double dx = 2.0 * xmax / (double)ncpx, dy = 2.0 * ymax / (double)ncpy,
               dz = 2.0 * zmax / (double)ncpz;
  const double ddx = dx / (double)nmesh, ddy = dy / (double)nmesh, ddz = dz / (double)nmesh;
  const double pi = acos(-1.0), fourpi = 4.0 * pi;
  const double omega = 8.0 * xmax * ymax * zmax, omegain = 1.0 / omega;
  const double surf = 4.0 * xmax * ymax;
  (void)ncpx_d;
#define CP(c, na) cp[(size_t)((c) - 1) + 8 * (size_t)(indspe[na] - 1)]

  for (int na = 0; na < natom; na++) {
    double tx = 0.0, ty = 0.0, tz = 0.0;
    force_01(na, nmesh, ncpx, ncpy, ncpz, num_list_d, key_natpri_inps, dx, dy, dz, &tx, &ty, &tz,
             natprid, napsd, natinfd, natinfd_vloc, lstvecd2, lstdx, lstdy, lstdz, rho_coarse,
             rho_aug_dense, vh_dense, drhoaug3ddx, drhoaug3ddy, drhoaug3ddz, dvlocdx_scw,
             dvlocdy_scw, dvlocdz_scw, dvlocdx_hdp, dvlocdy_hdp, dvlocdz_hdp, ndatx, ndaty, ndatz,
             biasx, biasy, biasz, vloc_dense);
    fatx[na] = fatx[na] + tx;
    faty[na] = faty[na] + ty;
    fatz[na] = fatz[na] + tz;
  }
  force_02(natom, nspv, nradmx, nqmx, ncpx * nmesh, ncpy * nmesh, ncpz * nmesh, xmax, ymax, zmax, ddx,
           ddy, ddz, psctoff, psftrad, nradct, indspe, nqctpcc, fatx, faty, fatz, atx, aty, atz,
           radial, coef, vx_dense);

  const size_t nd = (size_t)ncpx * nmesh * ncpy * nmesh * ncpz * nmesh;
  (void)nd;
  /* jellium part I (:1741-1774) */
  double fj = 0.0;
  if (jelcalc == key_jel_calc) {
    for (int na = 0; na < natom; na++) {
      const double cp1 = CP(1, na), cp2 = CP(2, na), cp3 = CP(3, na), cp4 = CP(4, na), cp5 = CP(5, na);
      for (int kz = -nzmax; kz <= nzmax; kz++) {
        if (kz * kz == 0) continue;
        const double vkz = pi / zmax * kz;
        const double vk2 = vkz * vkz;
        double fj0 = 0.0;
        for (int iz = 1; iz <= ncpz * nmesh; iz++)
          for (int iy = 1; iy <= ncpy * nmesh; iy++)
            for (int ix = 1; ix <= ncpx * nmesh; ix++) {
              double x = ix * ddx - 0.5 * ddx - xmax - atx[na];
              double y = iy * ddy - 0.5 * ddy - ymax - aty[na];
              double z = iz * ddz - 0.5 * ddz - zmax - atz[na];
              const double zz = iz * ddz - 0.5 * ddz - zmax;
              if (x > xmax) x = x - 2.0 * xmax;
              if (x <= -xmax) x = x + 2.0 * xmax;
              if (y > ymax) y = y - 2.0 * ymax;
              if (y <= -ymax) y = y + 2.0 * ymax;
              if (z > zmax) z = z - 2.0 * zmax;
              if (z <= -zmax) z = z + 2.0 * zmax;
              const double r = sqrt(x * x + y * y + z * z);
              fj0 = fj0 - 4.0 * pi / omega * chrjel / (strjel - endjel) / vk2 *
                              (cos(vkz * (strjel - zz)) - cos(vkz * (endjel - zz))) * cp1 *
                              (cp4 * pow(cp2 / pi, 1.5) * exp(-(cp2 * r * r)) +
                               cp5 * pow(cp3 / pi, 1.5) * exp(-(cp3 * r * r))) *
                              ddx * ddy * ddz;
            }
        fj = fj + fj0;
        fatz[na] = fatz[na] - fj0;
      }
    }
  }

  /* core-core interaction (:1777-1932) */
  for (int l0 = 0; l0 < natom; l0++) {
    double fxx = 0.0, fyy = 0.0, fzz = 0.0;
    const double x0 = atx[l0], y0 = aty[l0], z0 = atz[l0];
    for (int l1 = 0; l1 < natom; l1++) {
      const double x1 = atx[l1], y1 = aty[l1], z1 = atz[l1];
      const double c0 = CP(1, l0), c1 = CP(1, l1);
      switch (nperi) {
        case 0: {
          const double x = x0 - x1, y = y0 - y1, z = z0 - z1;
          const double r = sqrt(x * x + y * y + z * z);
          if (r > 1.0e-20) {
            const double r3 = r * r * r;
            fxx = fxx + c0 * c1 * x / r3;
            fyy = fyy + c0 * c1 * y / r3;
            fzz = fzz + c0 * c1 * z / r3;
          }
          break;
        }
        case 1: {
          for (int kx = -new_pwx; kx <= new_pwx; kx++) {
            if (kx * kx == 0) continue;
            const double vkx = pi / xmax * kx;
            const double ta = (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1);
            const double tb = vkx * vkx;
            const double dt = veta / nint1dmax;
            double vlo0 = 0.0, vlo1 = 0.0;
            for (int i = 1; i <= nint1dmax; i++) {
              const double t = i * dt;
              const double e = exp(-(ta * t * t) - 0.25 * tb / (t * t));
              vlo0 = vlo0 + e / t * dt;
              vlo1 = vlo1 + e * t * dt;
            }
            const double vcosi = -cos(vkx * (x0 - x1)) / xmax * 0.5;
            const double vsine = -sin(vkx * (x0 - x1)) / xmax * 0.5;
            fxx = fxx - 2.0 * vkx * c0 * c1 * vsine * vlo0;
            fyy = fyy - 4.0 * c0 * c1 * vcosi * vlo1 * (y0 - y1);
            fzz = fzz - 4.0 * c0 * c1 * vcosi * vlo1 * (z0 - z1);
          }
          const double ta = (y0 - y1) * (y0 - y1) + (z0 - z1) * (z0 - z1);
          double vlo0r;
          if (ta > 1.0e-16) vlo0r = -2.0 * (1.0 - exp(-(veta * veta * ta))) / (ta * 2.0 * xmax);
          else vlo0r = -(veta * veta / xmax);
          fyy = fyy - c0 * c1 * vlo0r * (y0 - y1);
          fzz = fzz - c0 * c1 * vlo0r * (z0 - z1);
          for (int kx = -new_pwx; kx <= new_pwx; kx++)
            ewald_real(c0 * c1, veta, pi, kx * 2.0 * xmax + x0 - x1, y0 - y1, z0 - z1, &fxx, &fyy, &fzz);
          break;
        }
        case 2: {
          for (int ky = -new_pwy; ky <= new_pwy; ky++)
            for (int kx = -new_pwx; kx <= new_pwx; kx++) {
              if (kx * kx + ky * ky == 0) continue;
              const double vkx = pi / xmax * kx, vky = pi / ymax * ky;
              const double vk2 = vkx * vkx + vky * vky;
              const double ta = z0 - z1;
              const double tb = sqrt(vk2);
              const double e1 = 1.0 - erf((tb - 2.0 * veta * veta * ta) / (2.0 * veta));
              const double e2 = 1.0 - erf((tb + 2.0 * veta * veta * ta) / (2.0 * veta));
              const double vep0 = -(pi / tb / surf * (e1 / exp(tb * ta) + e2 / exp(-(tb * ta))));
              const double vep1 = -(pi / surf * (e1 * exp(-(tb * ta)) - e2 * exp(tb * ta)));
              const double vlo0 = sin(vkx * (x0 - x1) + vky * (y0 - y1)) * vep0;
              const double vlo1 = cos(vkx * (x0 - x1) + vky * (y0 - y1)) * vep1;
              fxx = fxx - vkx * c0 * c1 * vlo0;
              fyy = fyy - vky * c0 * c1 * vlo0;
              fzz = fzz - c0 * c1 * vlo1;
            }
          {
            const double z = fabs(atz[l0] - atz[l1]);
            if (z > 1.0e-20) {
              const double vlo1 = -(2.0 * sqrt(pi) / surf * c0 * c1 * (atz[l0] - atz[l1]) / z);
              fzz = fzz - vlo1 * sqrt(pi) * erf(veta * z);
            }
          }
          for (int ky = -new_rsy; ky <= new_rsy; ky++)
            for (int kx = -new_rsx; kx <= new_rsx; kx++)
              ewald_real(c0 * c1, veta, pi, kx * 2.0 * xmax + x0 - atx[l1], ky * 2.0 * ymax + y0 - aty[l1],
                         z0 - atz[l1], &fxx, &fyy, &fzz);
          break;
        }
        case 3: {
          for (int kz = -new_pwz; kz <= new_pwz; kz++)
            for (int ky = -new_pwy; ky <= new_pwy; ky++)
              for (int kx = -new_pwx; kx <= new_pwx; kx++) {
                if (kx * kx + ky * ky + kz * kz == 0) continue;
                const double vkx = pi / xmax * kx, vky = pi / ymax * ky, vkz = pi / zmax * kz;
                const double vk2 = vkx * vkx + vky * vky + vkz * vkz;
                double vlo0;
                if (vk2 / (4.0 * (veta * veta)) < 100.0)
                  vlo0 = -(fourpi * omegain * sin(vkx * (x0 - x1) + vky * (y0 - y1) + vkz * (z0 - z1)) / vk2 *
                           exp(-(vk2 / (4.0 * (veta * veta)))));
                else
                  vlo0 = 0.0;
                fxx = fxx - c0 * c1 * vlo0 * vkx;
                fyy = fyy - c0 * c1 * vlo0 * vky;
                fzz = fzz - c0 * c1 * vlo0 * vkz;
              }
          for (int kz = -new_rsz; kz <= new_rsz; kz++)
            for (int ky = -new_rsy; ky <= new_rsy; ky++)
              for (int kx = -new_rsx; kx <= new_rsx; kx++)
                ewald_real(c0 * c1, veta, pi, kx * 2.0 * xmax + x0 - atx[l1], ky * 2.0 * ymax + y0 - aty[l1],
                           kz * 2.0 * zmax + z0 - atz[l1], &fxx, &fyy, &fzz);
          break;
        }
        default:
          break;
      }
    }
    fatx[l0] = fatx[l0] + fxx;
    faty[l0] = faty[l0] + fyy;
    fatz[l0] = fatz[l0] + fzz;
  }

  /* jellium part II (:1943-1987) */
  if (jelcalc == key_jel_calc) {
    for (int na = 0; na < natom; na++) {
      if (natpri[na] != key_natpri_in) continue;
      const int ipri = natpri_inf[na], sp = indspe[na];
      for (int kz = -nzmax; kz <= nzmax; kz++) {
        if (kz * kz == 0) continue;
        const double vkz = pi / zmax * kz;
        const double vk2 = vkz * vkz;
        for (int ns = 0; ns < (nspv < 2 ? nspv : 2); ns++)
          for (int il = 0; il < npoint; il++)
            for (int ir = 2; ir <= nradct[sp - 1]; ir++) {
              const size_t rr = (size_t)(ir - 1) + (size_t)nradmx * (sp - 1);
              const double z = point[(size_t)il + (size_t)npoint * 2] * radial[rr] + atz[na];
              const double r = radial[rr];
              const double dr = dradial[rr];
              const size_t ia = (size_t)(ir - 1) + (size_t)nradmx * ((size_t)il + (size_t)npoint * ((size_t)ns + (size_t)nspv * (ipri - 1)));
              const size_t ib = (size_t)(ir - 1) + (size_t)nradmx * ((size_t)il + (size_t)npoint * (ipri - 1));
              fj = fj + 4.0 * pi / omega * chrjel / (strjel - endjel) / vk2 *
                            (cos(vkz * (strjel - z)) - cos(vkz * (endjel - z))) *
                            (rhotrur[ia] - rhosmtr[ia] - rhoaugr[ib]) * r * r * dr * wt[il];
            }
      }
    }
    for (int kz = -nzmax; kz <= nzmax; kz++) {
      if (kz * kz == 0) continue;
      const double vkz = pi / zmax * kz;
      const double vk2 = vkz * vkz;
      double fj0 = 0.0;
      for (int iz = 1; iz <= ncpz * nmesh; iz++)
        for (int iy = 1; iy <= ncpy * nmesh; iy++)
          for (int ix = 1; ix <= ncpx * nmesh; ix++) {
            const double z = iz * ddz - 0.5 * ddz - zmax;
            const size_t p = (size_t)(ix - 1) + (size_t)ncpx * nmesh * ((iy - 1) + (size_t)ncpy * nmesh * (iz - 1));
            fj0 = fj0 + 4.0 * pi / omega * chrjel / (strjel - endjel) / vk2 *
                            (cos(vkz * (strjel - z)) - cos(vkz * (endjel - z))) * rho_aug_dense[p] * ddx *
                            ddy * ddz;
          }
      fj = fj + fj0;
    }
  }

  /* bias field and the net force (:1990-2007) */
  for (int na = 0; na < natom; na++) {
    fatx[na] = fatx[na] + biasx * CP(1, na);
    faty[na] = faty[na] + biasy * CP(1, na);
    fatz[na] = fatz[na] + biasz * CP(1, na);
  }
  double fxall = 0.0, fyall = 0.0, fzall = fj;
  for (int na = 0; na < natom; na++) {
    fxall = fxall + fatx[na];
    fyall = fyall + faty[na];
    fzall = fzall + fatz[na];
  }
  for (int na = 0; na < natom; na++) {
    fatx[na] = fatx[na] - fxall / (double)natom;
    faty[na] = faty[na] - fyall / (double)natom;
    fatz[na] = fatz[na] - fzall / (double)natom;
  }
  *fjel = fj;
#undef CP
  return 0;
}
