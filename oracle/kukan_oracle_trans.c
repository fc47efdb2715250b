/*
 * kukan_oracle_trans.c — FP64 CPU restatement of the potential-assembly grid transfers of the
 * RSPACE kukan v8f solver: the coarse <-> dense interpolation (interpolation8f.F90), its halo
 * handling on one rank (overlap_interpolation8f.F90), the transfers built on them
 * (trans8f.F90), the augmented dense density (scf_rhoaugdense8f.F90) and the V_eff sum of
 * scf_potentials (scf_potentials8f.F90:255-289).
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like kukan_oracle.c: loaded by the tests, never by
 * the product library.  Every routine follows the reference routine cited above it (file:line
 * relative to RSPACEes.v08.01/kukan/) on one MPI rank, with its loop and summation order.
 *
 * Pins (tests/test_oracle_trans.py): Lagrange interpolation of order 2 nf reproduces
 * polynomials of degree < 2 nf exactly; the restriction is the adjoint of the prolongation
 * divided by nmesh^3 (<R d, c> = <d, P c> / nmesh^3, periodic and isolated); the multipole
 * halo against the closed-form potential of point multipoles; the augmented density against
 * numpy's unbuffered scatter-add; and every routine against the compiled reference
 * (tests/test_reference_trans.py).
 */
#include "kukan_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int orc_set_error(int code, const char *msg); /* kukan_oracle.c */

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "kukan_oracle_trans: out of memory\n");
    abort();
  }
  return p;
}

/* index into an array with bounds lo..hi on each axis (column-major), extents e0, e1 */
#define A3(a, i, j, k, lo0, lo1, lo2, e0, e1) \
  (a)[(size_t)((i) - (lo0)) + (size_t)(e0) * ((size_t)((j) - (lo1)) + (size_t)(e1) * (size_t)((k) - (lo2)))]

/* ------------------------------------------------------------------ overlap_interpolation8f.F90:8-892
 * One rank.  nflag = 1: the halo of nf layers is filled by the periodic wrap along each
 * periodic axis (x for nperi >= 1, y for nperi >= 2, z for nperi = 3), over the full extent of
 * the other two axes, x then y then z; isolated axes keep their halo.  nflag = 2: along each
 * periodic axis the halo layers are added onto the interior layers they wrap to (captured
 * before the additions, :492-560), x then y then z; the halo of an isolated axis is dropped
 * (its additions exist only between ranks). */
int orc_overlap_interpolation(int ndisp, int nperi, int ncpa, int ncpb, int ncpc, int nf, int nf1,
                              double *v_cc, int nflag) {
  (void)ndisp;
  if (nflag != 1 && nflag != 2) return orc_set_error(1, "error in overlap_interpolation. nflag should be 1 or 2!");
  const int lo = -nf1;
  const long e0 = ncpa + nf + nf1 + 1, e1 = ncpb + nf + nf1 + 1, e2 = ncpc + nf + nf1 + 1;
#define V(i, j, k) A3(v_cc, i, j, k, lo, lo, lo, e0, e1)
  double *s2 = (double *)xcalloc((size_t)nf * (e0 > e1 ? e0 : e1) * (e1 > e2 ? e1 : e2) * 2 + 1, sizeof(double));
  for (int axis = 0; axis < 3; axis++) {
    if (nperi < axis + 1) continue;
    const int n = axis == 0 ? ncpa : axis == 1 ? ncpb : ncpc;
    /* the two other axes, full extent */
    const int m1lo = lo, m1hi = axis == 0 ? ncpb + nf : ncpa + nf;
    const int m2lo = lo, m2hi = axis == 2 ? ncpb + nf : ncpc + nf;
    const long c1 = m1hi - m1lo + 1;
    const long plane = c1 * (m2hi - m2lo + 1);
    double *lowc = s2, *highc = s2 + (size_t)nf * plane; /* captured halos */
#define AT(iov_idx, p, q) (axis == 0 ? &V(iov_idx, p, q) : axis == 1 ? &V(p, iov_idx, q) : &V(p, q, iov_idx))
    for (int q = m2lo; q <= m2hi; q++)
      for (int p = m1lo; p <= m1hi; p++)
        for (int iov = 1; iov <= nf; iov++) {
          const size_t t = (size_t)(iov - 1) + (size_t)nf * ((size_t)(p - m1lo) + (size_t)c1 * (q - m2lo));
          if (nflag == 1) {
            lowc[t] = *AT(n - nf + iov, p, q);
            highc[t] = *AT(iov, p, q);
          } else {
            lowc[t] = *AT(-nf + iov, p, q);
            highc[t] = *AT(n + iov, p, q);
          }
        }
    for (int q = m2lo; q <= m2hi; q++)
      for (int p = m1lo; p <= m1hi; p++)
        for (int iov = 1; iov <= nf; iov++) {
          const size_t t = (size_t)(iov - 1) + (size_t)nf * ((size_t)(p - m1lo) + (size_t)c1 * (q - m2lo));
          if (nflag == 1) {
            *AT(-nf + iov, p, q) = lowc[t];
            *AT(n + iov, p, q) = highc[t];
          } else {
            *AT(iov, p, q) = *AT(iov, p, q) + highc[t];
            *AT(n - nf + iov, p, q) = *AT(n - nf + iov, p, q) + lowc[t];
          }
        }
#undef AT
  }
  free(s2);
#undef V
  return 0;
}

/* Lagrange weights aweight(j, k), j = -nf1..nf, k = 0..nmesh-1 (interpolation8f.F90:40-50) */
static void lagrange_weights(int nf, int nmesh, double *w /* [2nf][nmesh] */) {
  const int nf1 = nf - 1;
  const double tmp = 1.0 / nmesh, tmp1 = (nmesh % 2 == 1) ? 0.0 : 0.5;
  for (int k = 0; k < nmesh; k++) {
    const double r = ((double)k + tmp1) * tmp;
    for (int j = -nf1; j <= nf; j++) {
      double a = 1.0;
      for (int i = -nf1; i <= nf; i++)
        if (i != j) a = a * (r - (double)i) / ((double)j - (double)i);
      w[(j + nf1) + 2 * nf * k] = a;
    }
  }
}
#define AW(j, k) aw[((j) + nf1) + 2 * nf * (k)]

/* ------------------------------------------------------------------ interpolation8f.F90:8-225
 * is = 1: v_dd (interior) = separable order-2nf Lagrange interpolation of the halo'd v_cc,
 * x then y then z.  is = 2: the transpose, summed onto the coarse halo'd grid z then y then x
 * from the whole dense array (its halo zeroed first along the periodic axes), divided by
 * nmesh^3.  One rank. */
int orc_interpolation(int ncpx, int ncpy, int ncpz, int nmesh, double *v_dd, double *v_cc, int nperi,
                      int nf, int is, int ndisp) {
  (void)ndisp;
  if (is != 1 && is != 2) return orc_set_error(1, "error in interpolation! IS should be 1 or 2!");
  const int nf1 = nf - 1, nf2 = nf * nmesh, nf3 = nf2 - 1;
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  double *aw = (double *)xcalloc((size_t)2 * nf * nmesh, sizeof(double));
  lagrange_weights(nf, nmesh, aw);
  /* extents */
  const long cx = ncpx + 2 * nf, cy = ncpy + 2 * nf, cz = ncpz + 2 * nf;
  const long dxe = nxd + 2 * nf2, dye = nyd + 2 * nf2, dze = nzd + 2 * nf2;
  /* vtmp1(-nf3:nxd+nf2, -nf1:ncpy+nf, -nf1:ncpz+nf), vtmp2(-nf3:nxd+nf2, -nf3:nyd+nf2, -nf1:ncpz+nf) */
  double *t1 = (double *)xcalloc((size_t)dxe * cy * cz, sizeof(double));
  double *t2 = (double *)xcalloc((size_t)dxe * dye * cz, sizeof(double));
#define VC(i, j, k) A3(v_cc, i, j, k, -nf1, -nf1, -nf1, cx, cy)
#define VD(i, j, k) A3(v_dd, i, j, k, -nf3, -nf3, -nf3, dxe, dye)
#define T1(i, j, k) A3(t1, i, j, k, -nf3, -nf1, -nf1, dxe, cy)
#define T2(i, j, k) A3(t2, i, j, k, -nf3, -nf3, -nf1, dxe, dye)
  const int hm = (nmesh - 1) / 2;
  if (is == 1) {
    memset(v_dd, 0, sizeof(double) * (size_t)dxe * dye * dze);
    for (int jz = -nf1; jz <= ncpz + nf; jz++)
      for (int jy = -nf1; jy <= ncpy + nf; jy++)
        for (int jx = 1; jx <= nxd; jx++)
          for (int ix = -nf1; ix <= nf; ix++) {
            const int kx = (jx + hm) % nmesh, lx = (jx + hm) / nmesh;
            T1(jx, jy, jz) = T1(jx, jy, jz) + AW(ix, kx) * VC(lx + ix, jy, jz);
          }
    for (int jz = -nf1; jz <= ncpz + nf; jz++)
      for (int jy = 1; jy <= nyd; jy++)
        for (int jx = 1; jx <= nxd; jx++)
          for (int iy = -nf1; iy <= nf; iy++) {
            const int ky = (jy + hm) % nmesh, ly = (jy + hm) / nmesh;
            T2(jx, jy, jz) = T2(jx, jy, jz) + AW(iy, ky) * T1(jx, ly + iy, jz);
          }
    for (int jz = 1; jz <= nzd; jz++)
      for (int jy = 1; jy <= nyd; jy++)
        for (int jx = 1; jx <= nxd; jx++)
          for (int iz = -nf1; iz <= nf; iz++) {
            const int kz = (jz + hm) % nmesh, lz = (jz + hm) / nmesh;
            VD(jx, jy, jz) = VD(jx, jy, jz) + AW(iz, kz) * T2(jx, jy, lz + iz);
          }
  } else {
    memset(v_cc, 0, sizeof(double) * (size_t)cx * cy * cz);
    /* the halo along the periodic axes is cleared (:126-177; one rank) */
    for (int jz = -nf3; jz <= nzd + nf2; jz++)
      for (int jy = -nf3; jy <= nyd + nf2; jy++)
        for (int jx = -nf3; jx <= nxd + nf2; jx++) {
          const int hx = jx < 1 || jx > nxd, hy = jy < 1 || jy > nyd, hz = jz < 1 || jz > nzd;
          if ((hx && nperi > 0) || (hy && nperi > 1) || (hz && nperi > 2)) VD(jx, jy, jz) = 0.0;
        }
    const int off = nf * nmesh + hm;
    for (int iz = -nf1; iz <= nf; iz++)
      for (int jz = -nf3; jz <= nzd + nf2; jz++) {
        const int lz = (jz + off) / nmesh - nf;
        if ((lz + iz + nf1) * (ncpz + nf - (lz + iz)) >= 0) {
          const int kz = (jz + off) % nmesh;
          for (int jy = -nf3; jy <= nyd + nf2; jy++)
            for (int jx = -nf3; jx <= nxd + nf2; jx++)
              T2(jx, jy, lz + iz) = T2(jx, jy, lz + iz) + AW(iz, kz) * VD(jx, jy, jz);
        }
      }
    for (int iy = -nf1; iy <= nf; iy++)
      for (int jz = -nf1; jz <= ncpz + nf; jz++)
        for (int jy = -nf3; jy <= nyd + nf2; jy++) {
          const int ly = (jy + off) / nmesh - nf;
          if ((ly + iy + nf1) * (ncpy + nf - (ly + iy)) >= 0) {
            const int ky = (jy + off) % nmesh;
            for (int jx = -nf3; jx <= nxd + nf2; jx++)
              T1(jx, ly + iy, jz) = T1(jx, ly + iy, jz) + AW(iy, ky) * T2(jx, jy, jz);
          }
        }
    for (int ix = -nf1; ix <= nf; ix++)
      for (int jz = -nf1; jz <= ncpz + nf; jz++)
        for (int jy = -nf1; jy <= ncpy + nf; jy++)
          for (int jx = -nf3; jx <= nxd + nf2; jx++) {
            const int lx = (jx + off) / nmesh - nf;
            if ((lx + ix + nf1) * (ncpx + nf - (lx + ix)) >= 0) {
              const int kx = (jx + off) % nmesh;
              VC(lx + ix, jy, jz) = VC(lx + ix, jy, jz) + AW(ix, kx) * T1(jx, jy, jz);
            }
          }
    const double n3 = (double)nmesh * nmesh * nmesh;
    for (size_t i = 0; i < (size_t)cx * cy * cz; i++) v_cc[i] = v_cc[i] / n3;
  }
#undef VC
#undef VD
#undef T1
#undef T2
  free(aw); free(t1); free(t2);
  return 0;
}

static double *halo_c(int ncpx, int ncpy, int ncpz, int nf) {
  return (double *)xcalloc((size_t)(ncpx + 2 * nf) * (ncpy + 2 * nf) * (ncpz + 2 * nf), sizeof(double));
}
static double *halo_d(int ncpx, int ncpy, int ncpz, int nf, int nmesh) {
  const int nf2 = nf * nmesh;
  return (double *)xcalloc((size_t)(ncpx * nmesh + 2 * nf2) * (ncpy * nmesh + 2 * nf2) * (ncpz * nmesh + 2 * nf2),
                           sizeof(double));
}

/* ------------------------------------------------------------------ trans8f.F90:8-47 */
int orc_trans_c2d_smtcharge(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                            const double *rhosmt, double *rho_dense) {
  const int nf1 = nf - 1, nf2 = nf * nmesh, nf3 = nf2 - 1;
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  const long cx = ncpx + 2 * nf, cy = ncpy + 2 * nf, dx = nxd + 2 * nf2, dy = nyd + 2 * nf2;
  const size_t n = (size_t)ncpx * ncpy * ncpz, nd = (size_t)nxd * nyd * nzd;
  double *vc = halo_c(ncpx, ncpy, ncpz, nf), *vd = halo_d(ncpx, ncpy, ncpz, nf, nmesh);
  int rc = 0;
  for (int ns = 0; ns < nspv && !rc; ns++) {
    for (int iz = 1; iz <= ncpz; iz++)
      for (int iy = 1; iy <= ncpy; iy++)
        for (int ix = 1; ix <= ncpx; ix++)
          A3(vc, ix, iy, iz, -nf1, -nf1, -nf1, cx, cy) =
              rhosmt[(size_t)(ix - 1) + ncpx * ((size_t)(iy - 1) + ncpy * (size_t)(iz - 1)) + n * ns];
    if ((rc = orc_overlap_interpolation(ndisp, nperi, ncpx, ncpy, ncpz, nf, nf1, vc, 1))) break;
    if ((rc = orc_interpolation(ncpx, ncpy, ncpz, nmesh, vd, vc, nperi, nf, 1, ndisp))) break;
    for (int iz = 1; iz <= nzd; iz++)
      for (int iy = 1; iy <= nyd; iy++)
        for (int ix = 1; ix <= nxd; ix++)
          rho_dense[(size_t)(ix - 1) + nxd * ((size_t)(iy - 1) + nyd * (size_t)(iz - 1)) + nd * ns] =
              A3(vd, ix, iy, iz, -nf3, -nf3, -nf3, dx, dy);
  }
  free(vc); free(vd);
  return rc;
}

/* dense -> coarse with the dense halo given (interpolation is = 2, overlap_interpolation 2) */
static int d2c(int nmesh, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz, double *vd,
               double *out) {
  const int nf1 = nf - 1;
  const long cx = ncpx + 2 * nf, cy = ncpy + 2 * nf;
  double *vc = halo_c(ncpx, ncpy, ncpz, nf);
  int rc = orc_interpolation(ncpx, ncpy, ncpz, nmesh, vd, vc, nperi, nf, 2, ndisp);
  if (!rc) rc = orc_overlap_interpolation(ndisp, nperi, ncpx, ncpy, ncpz, nf, nf1, vc, 2);
  if (!rc)
    for (int iz = 1; iz <= ncpz; iz++)
      for (int iy = 1; iy <= ncpy; iy++)
        for (int ix = 1; ix <= ncpx; ix++)
          out[(size_t)(ix - 1) + ncpx * ((size_t)(iy - 1) + ncpy * (size_t)(iz - 1))] =
              A3(vc, ix, iy, iz, -nf1, -nf1, -nf1, cx, cy);
  free(vc);
  return rc;
}

static void load_dense(int nmesh, int nf, int ncpx, int ncpy, int ncpz, const double *src, double *vd) {
  const int nf2 = nf * nmesh, nf3 = nf2 - 1;
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  const long dx = nxd + 2 * nf2, dy = nyd + 2 * nf2;
  for (int iz = 1; iz <= nzd; iz++)
    for (int iy = 1; iy <= nyd; iy++)
      for (int ix = 1; ix <= nxd; ix++)
        A3(vd, ix, iy, iz, -nf3, -nf3, -nf3, dx, dy) = src[(size_t)(ix - 1) + nxd * ((size_t)(iy - 1) + nyd * (size_t)(iz - 1))];
}

/* ------------------------------------------------------------------ trans8f.F90:50-90 */
int orc_trans_d2c_smtcharge(int nmesh, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                            const double *rho_aug_dense, double *rho_coarse) {
  double *vd = halo_d(ncpx, ncpy, ncpz, nf, nmesh);
  load_dense(nmesh, nf, ncpx, ncpy, ncpz, rho_aug_dense, vd);
  const int rc = d2c(nmesh, nperi, ndisp, nf, ncpx, ncpy, ncpz, vd, rho_coarse);
  free(vd);
  return rc;
}

/* potential of the atom's point multipoles added on one dense halo point
 * (trans8f.F90:174-246, the expression of each face loop) */
static double multipole_term(double v, int jx, int jy, int jz, double ddx, double ddy, double ddz,
                             double xmax, double ymax, double zmax, const double *mp, double cp1,
                             double ax, double ay, double az) {
  const double atptmp = mp[0] / cp1, edpolx = mp[1] / cp1, edpoly = mp[2] / cp1, edpolz = mp[3] / cp1;
  const double eqpoxx = mp[4] / cp1, eqpoyy = mp[5] / cp1, eqpozz = mp[6] / cp1;
  const double eqpoxy = mp[7] / cp1, eqpoyz = mp[8] / cp1, eqpozx = mp[9] / cp1;
  const double eqpoxy2 = 2.0 * eqpoxy, eqpoyz2 = 2.0 * eqpoyz, eqpozx2 = 2.0 * eqpozx;
  const double x = ax - (jx * ddx - xmax - 0.5 * ddx);
  const double y = ay - (jy * ddy - ymax - 0.5 * ddy);
  const double z = az - (jz * ddz - zmax - 0.5 * ddz);
  const double r = sqrt(x * x + y * y + z * z);
  const double rin = 1.0 / r, r2in = rin * rin, r3in = r2in * rin, r5in = r3in * r2in;
  const double vmopm = -cp1 * rin, derfdrr = -r3in;
  const double vdpmx = cp1 * derfdrr * x, vdpmy = cp1 * derfdrr * y, vdpmz = cp1 * derfdrr * z;
  const double vqmxx = -cp1 * (3.0 * x * x * r5in - r3in), vqmyy = -cp1 * (3.0 * y * y * r5in - r3in);
  const double vqmzz = -cp1 * (3.0 * z * z * r5in - r3in), vqmxy = -cp1 * 3.0 * x * y * r5in;
  const double vqmyz = -cp1 * 3.0 * y * z * r5in, vqmzx = -cp1 * 3.0 * z * x * r5in;
  return v - atptmp * vmopm - edpolx * vdpmx - edpoly * vdpmy - edpolz * vdpmz - eqpoxx * vqmxx -
         eqpoyy * vqmyy - eqpozz * vqmzz - eqpoxy2 * vqmxy - eqpoyz2 * vqmyz - eqpozx2 * vqmzx;
}

/* ------------------------------------------------------------------ trans8f.F90:155-412 (one rank) */
static void vh_multipole_halo(int natom, int nmesh, int nperi, int nf, int nxd, int nyd, int nzd,
                              double ddx, double ddy, double ddz, double xmax, double ymax, double zmax,
                              const int *indspe, const double *atmpole, const double *cp,
                              const double *atx, const double *aty, const double *atz, double *vd) {
  const int nf2 = nf * nmesh, nf3 = nf2 - 1;
  const long dx = nxd + 2 * nf2, dy = nyd + 2 * nf2;
#define VD(i, j, k) A3(vd, i, j, k, -nf3, -nf3, -nf3, dx, dy)
#define TERM(i, j, k) VD(i, j, k) = multipole_term(VD(i, j, k), i, j, k, ddx, ddy, ddz, xmax, ymax, zmax, mp, cp1, atx[na], aty[na], atz[na])
  for (int na = 0; na < natom; na++) {
    const double cp1 = cp[8 * (size_t)(indspe[na] - 1)];
    const double *mp = atmpole + 10 * (size_t)na;
    if (nperi == 0) {
      for (int iz = -nf3; iz <= nzd + nf2; iz++)
        for (int iy = -nf3; iy <= nyd + nf2; iy++)
          for (int ix = -nf3; ix <= 0; ix++) TERM(ix, iy, iz);
      for (int iz = -nf3; iz <= nzd + nf2; iz++)
        for (int iy = -nf3; iy <= nyd + nf2; iy++)
          for (int ix = nxd + 1; ix <= nxd + nf2; ix++) TERM(ix, iy, iz);
    }
    if (nperi <= 1) {
      for (int iz = -nf3; iz <= nzd + nf2; iz++)
        for (int iy = -nf3; iy <= 0; iy++)
          for (int ix = -nf3; ix <= nxd + nf2; ix++) TERM(ix, iy, iz);
      for (int iz = -nf3; iz <= nzd + nf2; iz++)
        for (int iy = nyd + 1; iy <= nyd + nf2; iy++)
          for (int ix = -nf3; ix <= nxd + nf2; ix++) TERM(ix, iy, iz);
    }
    if (nperi <= 2) {
      for (int iz = -nf3; iz <= 0; iz++)
        for (int iy = -nf3; iy <= nyd + nf2; iy++)
          for (int ix = -nf3; ix <= nxd + nf2; ix++) TERM(ix, iy, iz);
      for (int iz = nzd + 1; iz <= nzd + nf2; iz++)
        for (int iy = -nf3; iy <= nyd + nf2; iy++)
          for (int ix = -nf3; ix <= nxd + nf2; ix++) TERM(ix, iy, iz);
    }
  }
#undef TERM
#undef VD
}

/* ------------------------------------------------------------------ trans8f.F90:93-152 */
int orc_trans_d2c_vh(int natom, int num_spe, int nmesh, int nperi, int ndisp, int nf, int ncpx,
                     int ncpy, int ncpz, double xmax, double ymax, double zmax, const int *indspe,
                     const double *atmpole, const double *cp, const double *atx, const double *aty,
                     const double *atz, double *vh_coarse, const double *vh_dense) {
  (void)num_spe;
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  const double ddx = 2.0 * xmax / nxd, ddy = 2.0 * ymax / nyd, ddz = 2.0 * zmax / nzd;
  double *vd = halo_d(ncpx, ncpy, ncpz, nf, nmesh);
  load_dense(nmesh, nf, ncpx, ncpy, ncpz, vh_dense, vd);
  if (nperi < 3)
    vh_multipole_halo(natom, nmesh, nperi, nf, nxd, nyd, nzd, ddx, ddy, ddz, xmax, ymax, zmax, indspe,
                      atmpole, cp, atx, aty, atz, vd);
  const int rc = d2c(nmesh, nperi, ndisp, nf, ncpx, ncpy, ncpz, vd, vh_coarse);
  free(vd);
  return rc;
}

/* ------------------------------------------------------------------ trans8f.F90:415-508
 * The isolated faces of the dense halo hold a constant (trans8f.F90:429, the reference's value
 * of the exchange-correlation potential far from the system), set once before the spins. */
int orc_trans_d2c_vxc(int nmesh, int nspv, int nperi, int ndisp, int nf, int ncpx, int ncpy, int ncpz,
                      double *vx, const double *vx_dense) {
  const int nf2 = nf * nmesh, nf3 = nf2 - 1;
  const int nxd = ncpx * nmesh, nyd = ncpy * nmesh, nzd = ncpz * nmesh;
  const long dx = nxd + 2 * nf2, dy = nyd + 2 * nf2;
  const double tmp1 = -1.865282921333129e-005;
  double *vd = halo_d(ncpx, ncpy, ncpz, nf, nmesh);
  for (int iz = -nf3; iz <= nzd + nf2; iz++)
    for (int iy = -nf3; iy <= nyd + nf2; iy++)
      for (int ix = -nf3; ix <= nxd + nf2; ix++) {
        const int hx = ix < 1 || ix > nxd, hy = iy < 1 || iy > nyd, hz = iz < 1 || iz > nzd;
        if ((hx && nperi == 0) || (hy && nperi <= 1) || (hz && nperi <= 2))
          A3(vd, ix, iy, iz, -nf3, -nf3, -nf3, dx, dy) = tmp1;
      }
  const size_t n = (size_t)ncpx * ncpy * ncpz, nd = (size_t)nxd * nyd * nzd;
  int rc = 0;
  for (int ns = 0; ns < nspv && !rc; ns++) {
    load_dense(nmesh, nf, ncpx, ncpy, ncpz, vx_dense + nd * ns, vd);
    rc = d2c(nmesh, nperi, ndisp, nf, ncpx, ncpy, ncpz, vd, vx + n * ns);
  }
  free(vd);
  return rc;
}

/* ------------------------------------------------------------------ scf_rhoaugdense8f.F90:8-48 */
int orc_scf_rhoaugdense(int key_natpri_inps, int key_pp_paw, int natom, int num_spe, int nspv,
                        int ncpx_d, int ncpy_d, int ncpz_d, int num_list_d, int num_ppcell_d,
                        const int *indspe, const int *natprid, const int *napsd, const int *natinfd,
                        const int *ntyppp, const int *lstvecd2, const double *rho_dense,
                        const double *rhoaug3d, double *rho_aug_dense) {
  (void)num_spe; (void)num_ppcell_d;
  const size_t nd = (size_t)ncpx_d * ncpy_d * ncpz_d;
  for (size_t i = 0; i < nd; i++) rho_aug_dense[i] = rho_dense[i];
  if (nspv > 1)
    for (size_t i = 0; i < nd; i++) rho_aug_dense[i] = rho_aug_dense[i] + rho_dense[i + nd];
  for (int na = 0; na < natom; na++) {
    if (ntyppp[indspe[na] - 1] != key_pp_paw || natprid[na] != key_natpri_inps) continue;
    const size_t ia = (size_t)(napsd[na] - 1) * num_list_d;
    for (int i = 0; i < natinfd[na]; i++) {
      const int p = lstvecd2[ia + i] - 1;
      rho_aug_dense[p] = rho_aug_dense[p] + rhoaug3d[ia + i];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ scf_potentials8f.F90:255-289
 * veff = vx + (vloc + vh + bias . r) for the first two spin channels (one rank). */
int orc_scf_veff(int nspv, int ncpx, int ncpy, int ncpz, double xmax, double ymax, double zmax,
                 double dx, double dy, double dz, double biasx, double biasy, double biasz,
                 const double *vloc_coarse, const double *vh_coarse, const double *vx, double *veff) {
  const size_t n = (size_t)ncpx * ncpy * ncpz;
  for (int iz = 1; iz <= ncpz; iz++)
    for (int iy = 1; iy <= ncpy; iy++)
      for (int ix = 1; ix <= ncpx; ix++) {
        const size_t p = (size_t)(ix - 1) + ncpx * ((size_t)(iy - 1) + ncpy * (size_t)(iz - 1));
        const double vdia = vloc_coarse[p] + vh_coarse[p] + biasx * (((double)ix - 0.5) * dx - xmax) +
                            biasy * (((double)iy - 0.5) * dy - ymax) + biasz * (((double)iz - 0.5) * dz - zmax);
        for (int ns = 0; ns < nspv; ns++) {
          veff[p + n * ns] = vx[p + n * ns];
          if (nspv < 3 || ns < 2) veff[p + n * ns] = veff[p + n * ns] + vdia;
        }
      }
  return 0;
}
